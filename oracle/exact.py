"""Oracle: exact / reference solutions used for L1 errors (numpy float64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

  * ho_exact       HarmonicOscillator1D.exact_solution, problems.py:134-146
  * poisson_exact  closed form of u_xx+u_yy = sum_k sin(2^k pi x) sin(2^k pi y)
                   with zero Dirichlet BC: u = -sum_k sin(.)sin(.)/(2 (2^k pi)^2)
  * burgers_exact  burgers_viscous_time_exact1, burgers_solution.py:40-144,
                   Gauss-Hermite rule of order 60 (burgers_solution.py:104,
                   hermite_ek_compute :479) == numpy hermgauss(60)

Pinned by tests/test_oracle_exact.py: the HO solution satisfies the ODE and
u(0)=1, u'(0)=0 (finite differences); the Poisson solution satisfies the PDE
(finite differences) and the BCs; the Burgers solution equals -sin(pi x) at
t=0, vanishes at x=+-1 and has a small FD residual away from the shock.
"""

import numpy as np


def ho_exact(x, d, w0):
    w = np.sqrt(w0 ** 2 - d ** 2)
    phi = np.arctan(-d / w)
    A = 1 / (2 * np.cos(phi))
    return np.exp(-d * x) * 2 * A * np.cos(phi + w * x)


def poisson_exact(x, ks=(1, 2, 3, 4)):
    x = np.asarray(x, dtype=np.float64)
    u = np.zeros((x.shape[0], 1))
    for k in ks:
        a = 2 ** k * np.pi
        u[:, 0] -= np.sin(a * x[:, 0]) * np.sin(a * x[:, 1]) / (2 * a * a)
    return u


def burgers_exact(nu, vx, vt, qn=60):
    """Returns vu (len(vx), len(vt))."""
    qx, qw = np.polynomial.hermite.hermgauss(qn)
    vx = np.asarray(vx, dtype=np.float64)
    vt = np.asarray(vt, dtype=np.float64)
    vu = np.zeros((len(vx), len(vt)))
    for j, t in enumerate(vt):
        if t == 0.0:
            vu[:, j] = -np.sin(np.pi * vx)
            continue
        c = 2.0 * np.sqrt(nu * t)
        arg = vx[:, None] - c * qx[None, :]
        e = np.exp(-np.cos(np.pi * arg) / (2.0 * np.pi * nu))
        top = -(qw[None, :] * c * np.sin(np.pi * arg) * e).sum(1)
        bot = (qw[None, :] * c * e).sum(1)
        vu[:, j] = top / bot
    return vu
