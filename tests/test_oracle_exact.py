"""Pins oracle/exact.py and checks the product's reference fields against it.

oracle/exact.py is pinned by what the mathematics fixes:
  * ho_exact: the ODE u'' + 2d u' + w0^2 u = 0 by central differences and
    u(0) = 1, u'(0) = 0;
  * poisson_exact: the PDE by central differences and zero boundary values;
  * burgers_exact: u(x, 0) = -sin(pi x), u(+-1, t) = 0 and the PDE residual by
    central differences away from the steepening front;
  * wave_free_space: the closed form at a source centre (Dawson's integral),
    the initial Gaussians at t = 0, u_t(., 0) = 0 and the PDE residual by
    central differences in (x, y, t).
The product's fields (fbpinns_amd/traditional_solutions/*, used by the
problems' exact_solution for the L1 metric) are then checked against the
pinned oracle: the Burgers quadrature to 1e-12, the C-PML finite-difference
solver by its second-order convergence to the free-space field (shifted by the
half time step of its u_t(0) = 0 start, p^-1 = p^0), its absorbing frame, and
the problem's refined-grid/decimation wrapper on a coarse test grid.
"""

import numpy as np
import pytest
import torch

from oracle import exact as oex
from oracle import problems as opb


def _fd2(f, x, h):
    return (f(x + h) - 2 * f(x) + f(x - h)) / (h * h), (f(x + h) - f(x - h)) / (2 * h)


def test_ho_exact_satisfies_ode_and_ics():
    d, w0 = 2, 80
    f = lambda t: oex.ho_exact(t, d, w0)
    t = np.linspace(0.05, 0.95, 50)
    utt, ut = _fd2(f, t, 1e-5)
    res = utt + 2 * d * ut + w0 ** 2 * f(t)
    assert np.abs(res).max() < 1e-3 * w0 ** 2
    assert abs(f(np.array([0.0]))[0] - 1.0) < 1e-14
    assert abs((f(np.array([1e-7])) - f(np.array([-1e-7])))[0] / 2e-7) < 1e-5


def test_poisson_exact_solves_pde():
    r = np.random.default_rng(0)
    x = r.uniform(0.05, 0.95, (100, 2))
    h = 1e-4
    lap = np.zeros(100)
    for i in range(2):
        e = np.zeros(2)
        e[i] = h
        lap += ((oex.poisson_exact(x + e) - 2 * oex.poisson_exact(x) + oex.poisson_exact(x - e)) / h ** 2)[:, 0]
    f = opb.poisson_source(torch.tensor(x)).numpy()[:, 0]
    assert np.abs(lap - f).max() < 1e-4
    xb = np.array([[0.0, 0.3], [1.0, 0.6], [0.2, 0.0], [0.8, 1.0]])
    assert np.abs(oex.poisson_exact(xb)).max() < 1e-15


def test_burgers_exact_initial_boundary_and_residual():
    nu = 0.01 / np.pi
    vx = np.linspace(-1, 1, 201)
    u0 = oex.burgers_exact(nu, vx, [0.0])[:, 0]
    assert np.allclose(u0, -np.sin(np.pi * vx))
    ub = oex.burgers_exact(nu, np.array([-1.0, 1.0]), [0.2, 0.5])
    assert np.abs(ub).max() < 1e-12
    # PDE residual by central differences away from the steepening front at x = 0
    xs = np.array([-0.8, -0.6, 0.6, 0.8])
    t, h, k = 0.2, 1e-4, 1e-4
    u = lambda xx, tt: oex.burgers_exact(nu, xx, [tt])[:, 0]
    ut = (u(xs, t + k) - u(xs, t - k)) / (2 * k)
    ux = (u(xs + h, t) - u(xs - h, t)) / (2 * h)
    uxx = (u(xs + h, t) - 2 * u(xs, t) + u(xs - h, t)) / h ** 2
    res = ut + u(xs, t) * ux - nu * uxx
    assert np.abs(res).max() < 1e-5


# ---------------------------------------------------------------- free-space wave field

SRC2 = np.array([[0.1, -0.2, 0.2, 1.0], [-0.3, 0.25, 0.1, 0.5]])


def test_wave_free_space_centre_closed_form():
    t = np.linspace(0.0, 1.0, 21)
    for s, a in ((0.03, 1.0), (0.1, 0.7), (0.3, -2.0)):
        src = np.array([[0.2, -0.4, s, a]])
        u = oex.wave_free_space(np.tile(src[:, :2], (len(t), 1)), t, src)
        assert np.abs(u - oex.wave_free_space_centre(t, s, a)).max() < 1e-12
    # the closed form itself: 1 at t = 0, and the 2D decay ~ -s^2/(c t)^2 for c t >> s
    assert oex.wave_free_space_centre(0.0, 0.1) == 1.0
    big = oex.wave_free_space_centre(50.0, 1.0)
    assert abs(big + 1 / 50.0 ** 2) < 1e-5


def test_wave_free_space_initial_condition():
    r = np.random.default_rng(0)
    xy = r.uniform(-1, 1, (500, 2))
    u0 = oex.wave_free_space(xy, np.zeros(500), SRC2)
    g = sum(a * np.exp(-0.5 * ((xy - [cx, cy]) ** 2).sum(1) / s ** 2) for cx, cy, s, a in SRC2)
    assert np.abs(u0 - g).max() < 1e-13
    # u_t(., 0) = 0: one-sided difference is O(h) * u_tt
    h = 1e-6
    u1 = oex.wave_free_space(xy, np.full(500, h), SRC2)
    assert np.abs(u1 - u0).max() / h < 1e-3


@pytest.mark.parametrize("c0", [1.0, 1.7])
def test_wave_free_space_solves_pde(c0):
    r = np.random.default_rng(1)
    n = 200
    xy = r.uniform(-0.8, 0.8, (n, 2))
    t = r.uniform(0.1, 0.8, n)
    h = 1e-3
    f = lambda dx, dy, dt: oex.wave_free_space(xy + [dx, dy], t + dt, SRC2, c0=c0)
    u0 = f(0, 0, 0)
    uxx = (f(h, 0, 0) - 2 * u0 + f(-h, 0, 0)) / h ** 2
    uyy = (f(0, h, 0) - 2 * u0 + f(0, -h, 0)) / h ** 2
    utt = (f(0, 0, h) - 2 * u0 + f(0, 0, -h)) / h ** 2
    res = uxx + uyy - utt / c0 ** 2
    scale = np.abs(utt).max()
    assert np.abs(res).max() < 1e-4 * scale
    # the check is sensitive to the wave speed: the same field fails with another c
    assert np.abs(uxx + uyy - utt / (1.05 * c0) ** 2).max() > 1e-2 * scale


# ---------------------------------------------------------------- product fields vs the oracle

def test_product_burgers_quadrature_matches_oracle():
    from fbpinns_amd.traditional_solutions.burgers import burgers_viscous_time_exact1
    nu = 0.01 / np.pi
    vx = np.linspace(-1, 1, 101)
    vt = np.linspace(0, 1, 21)
    a = burgers_viscous_time_exact1(nu, len(vx), vx, len(vt), vt)
    b = oex.burgers_exact(nu, vx, vt)
    assert np.abs(a - b).max() < 1e-12


def _cpml_run(src_sigma, nref, T):
    """The product's C-PML solver on [-1, 1]^2 for one centred Gaussian, grid
    spacing sigma / nref, returned on the sigma-spaced subgrid at time T."""
    from fbpinns_amd.traditional_solutions.seismic_cpml import seismicCPML2D
    dx = src_sigma / nref
    n = int(round(2 / dx)) + 1
    dt = dx / (4 * np.sqrt(2))
    nst = int(np.ceil(T / dt))
    dt = T / nst
    xs = np.linspace(-1, 1, n)
    xx, yy = np.meshgrid(xs, xs, indexing="ij")
    p0 = np.exp(-0.5 * (xx ** 2 + yy ** 2) / src_sigma ** 2)
    npml = 10
    p0p = np.pad(p0, npml, mode="edge")
    c = np.ones_like(p0p)
    wf, _ = seismicCPML2D(n + 2 * npml, n + 2 * npml, nst + 1, dx, dx, dt, npml, c, np.ones_like(c),
                          (p0p.copy(), p0p.copy()), 1.0 / src_sigma, np.float64)
    sel = (slice(None, None, nref), slice(None, None, nref))
    u = wf[-1, npml:-npml, npml:-npml][sel]
    g = np.stack([xx[sel].ravel(), yy[sel].ravel()], 1)
    return u.ravel(), g, dt


def test_product_cpml_converges_to_free_space_at_second_order():
    """Before the wave reaches the frame: the error against the free-space
    field at T + dt/2 falls 4x per halving of the grid spacing."""
    src = np.array([[0.0, 0.0, 0.1, 1.0]])
    errs = []
    for nref in (5, 10):
        u, g, dt = _cpml_run(0.1, nref, 0.6)
        uo = oex.wave_free_space(g, np.full(len(g), 0.6 + 0.5 * dt), src)
        errs.append(np.abs(u - uo).max())
    assert errs[1] < 1e-3
    assert 3.5 < errs[0] / errs[1] < 4.5, errs


def test_product_cpml_frame_absorbs():
    """After the front has left [-1, 1]^2 (t = 1.5) the field still matches
    free space to 1e-2 of the initial amplitude: a reflecting edge would
    return the outgoing ring (amplitude ~0.08 at the edge)."""
    src = np.array([[0.0, 0.0, 0.1, 1.0]])
    u, g, dt = _cpml_run(0.1, 5, 1.5)
    uo = oex.wave_free_space(g, np.full(len(g), 1.5 + 0.5 * dt), src)
    assert np.abs(u - uo).max() < 1e-2


def test_product_wave_exact_solution_matches_free_space():
    """The problem's exact_solution (refined C-PML grid, cropped frame,
    decimation onto the test grid; problems.py:407-474) against the oracle
    on a 41 x 41 x 6 grid with sources of width 0.2 and 0.1."""
    from fbpinns_amd.problems import WaveEquationConstantVelocity3D as W
    sp, _ = W.init_params(c0=1, source=SRC2)
    shape = (41, 41, 6)
    xs = [np.linspace(-1, 1, shape[0]), np.linspace(-1, 1, shape[1]), np.linspace(0, 1, shape[2])]
    g = np.stack(np.meshgrid(*xs, indexing="ij"), -1).reshape(-1, 3)
    u = W.exact_solution({"static": {"problem": sp}}, torch.tensor(g), shape).numpy()[:, 0]
    uo = oex.wave_free_space(g[:, :2], g[:, 2], SRC2)
    assert np.abs(u - uo).max() < 1e-2
    assert np.abs(u - uo).mean() < 1e-2 * np.abs(uo).mean()
    t0 = g[:, 2] == 0
    assert np.abs(u[t0] - uo[t0]).max() < 1e-6
