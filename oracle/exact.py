"""Oracle: exact / reference solutions used for L1 errors (numpy float64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

  * ho_exact       HarmonicOscillator1D.exact_solution, problems.py:134-146
  * poisson_exact  closed form of u_xx+u_yy = sum_k sin(2^k pi x) sin(2^k pi y)
                   with zero Dirichlet BC: u = -sum_k sin(.)sin(.)/(2 (2^k pi)^2)
  * burgers_exact  burgers_viscous_time_exact1, burgers_solution.py:40-144,
                   Gauss-Hermite rule of order 60 (burgers_solution.py:104,
                   hermite_ek_compute :479) == numpy hermgauss(60)
  * wave_free_space free-space field of the (2+1)D wave problem's initial
                   condition (problems.py:346-405) by Hankel-transform
                   quadrature: the field that the reference's C-PML
                   finite-difference solution (problems.py:407-474,
                   seismic_CPML_2D_pressure_second_order.py:12-150)
                   approximates inside its absorbing frame
  * wave_free_space_centre  its closed form at a source centre

Pinned by tests/test_oracle_exact.py: the HO solution satisfies the ODE and
u(0)=1, u'(0)=0 (finite differences); the Poisson solution satisfies the PDE
(finite differences) and the BCs; the Burgers solution equals -sin(pi x) at
t=0, vanishes at x=+-1 and has a small FD residual away from the shock; the
free-space wave field equals the Dawson-integral closed form at the source
centres, the Gaussians at t=0, has u_t(.,0)=0 and satisfies the PDE by
finite differences (and fails it for a 5 % different wave speed).
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


def wave_free_space(xy, t, source, c0=1.0, q_max=9.0, nodes_per_panel=16):
    """Free-space solution of u_tt = c0^2 (u_xx + u_yy) on R^2 with
    u(x, y, 0) = sum_k a_k exp(-|xy - c_k|^2 / (2 s_k^2)) and u_t(x, y, 0) = 0,
    source rows [cx, cy, s, a] (the initial condition of problems.py:383-392).

    Each Gaussian is radially symmetric, so its 2D Fourier transform is the
    Hankel transform a s^2 exp(-s^2 k^2 / 2) and every mode evolves as
    cos(c0 k t):
        u_k(r, t) = a * int_0^inf exp(-q^2/2) cos(q c0 t / s) J0(q r / s) q dq,
    integrated over q in [0, q_max] (the factor exp(-q_max^2/2) < 3e-18 bounds
    the cut tail) by composite Gauss-Legendre panels, one panel per ~period of
    the fastest oscillation.  This is the field the C-PML finite-difference
    solver of the wave problem approximates inside its absorbing frame
    (problems.py:407-474); float64.  xy: (n, 2), t: (n,) -> u (n,)."""
    from scipy.special import j0
    xy = np.asarray(xy, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64).reshape(-1)
    source = np.asarray(source, dtype=np.float64)
    gx, gw = np.polynomial.legendre.leggauss(nodes_per_panel)
    u = np.zeros(xy.shape[0])
    for cx, cy, s, a in source:
        r = np.hypot(xy[:, 0] - cx, xy[:, 1] - cy) / s
        tau = c0 * t / s
        n_pan = int(np.ceil(q_max * (tau.max() + r.max()) / (2 * np.pi))) + 8
        edges = np.linspace(0.0, q_max, n_pan + 1)
        h = np.diff(edges)
        q = (edges[:-1, None] + 0.5 * h[:, None] * (gx[None] + 1)).reshape(-1)
        w = (0.5 * h[:, None] * gw[None]).reshape(-1) * q * np.exp(-0.5 * q * q)
        for lo in range(0, xy.shape[0], 2048):
            sl = slice(lo, lo + 2048)
            f = np.cos(tau[sl, None] * q[None]) * j0(r[sl, None] * q[None])
            u[sl] += a * (f @ w)
    return u


def wave_free_space_centre(t, s, a=1.0, c0=1.0):
    """Closed form of wave_free_space at a source centre (r = 0):
    int_0^inf q exp(-q^2/2) cos(q tau) dq = 1 - sqrt(2) tau D(tau / sqrt(2)),
    D = Dawson's integral, tau = c0 t / s."""
    from scipy.special import dawsn
    tau = c0 * np.asarray(t, dtype=np.float64) / s
    return a * (1.0 - np.sqrt(2.0) * tau * dawsn(tau / np.sqrt(2.0)))
