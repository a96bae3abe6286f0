"""Exact viscous Burgers solution by Gauss-Hermite quadrature.

Restates burgers_viscous_time_exact1 (traditional_solutions/analytical/
burgers_solution.py:40-144): for u_t + u u_x = nu u_xx on (-1, 1),
u(x, 0) = -sin(pi x), u(+-1, t) = 0, with c = 2 sqrt(nu t),
    u(x, t) = - sum_q w_q c sin(pi (x - c y_q)) e(x - c y_q)
              / sum_q w_q c e(x - c y_q),   e(s) = exp(-cos(pi s) / (2 pi nu)),
over the order-60 Gauss-Hermite rule (weight exp(-y^2)) used there
(:104, hermite_ek_compute :479) -- numpy's hermgauss gives the same rule.
Vectorised over x; float64.
"""

import numpy as np


def burgers_viscous_time_exact1(nu, vxn, vx, vtn, vt, qn=60):
    qx, qw = np.polynomial.hermite.hermgauss(qn)
    vx = np.asarray(vx, dtype=np.float64)[:vxn]
    vt = np.asarray(vt, dtype=np.float64)[:vtn]
    vu = np.zeros((vxn, vtn))
    for j, t in enumerate(vt):
        if t == 0.0:
            vu[:, j] = -np.sin(np.pi * vx)
            continue
        c = 2.0 * np.sqrt(nu * t)
        s = vx[:, None] - c * qx[None, :]
        e = np.exp(-np.cos(np.pi * s) / (2.0 * np.pi * nu))
        top = -(qw * c * np.sin(np.pi * s) * e).sum(1)
        bot = (qw * c * e).sum(1)
        vu[:, j] = top / bot
    return vu
