"""Pins oracle/adam.py (Kingma & Ba closed forms and an independent scalar
loop), oracle/problems.py (residuals vanish on exact solutions, hard BCs hold)
and oracle/exact.py (ODE/PDE residuals by finite differences, boundary and
initial values)."""

import math

import numpy as np
import torch

from oracle import adam as oadam
from oracle import exact as oex
from oracle import problems as opb


def test_adam_first_step_is_signed_lr():
    r = np.random.default_rng(0)
    p = r.standard_normal(50)
    g = r.standard_normal(50)
    p1, m, v, c = oadam.adam_update(p, np.zeros(50), np.zeros(50), g, 0, lr=1e-3, eps=1e-8)
    assert c == 1
    assert np.allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)


def test_adam_constant_gradient_closed_form():
    g = np.array([0.5, -2.0, 1e-3])
    p = np.zeros(3)
    m = np.zeros(3)
    v = np.zeros(3)
    c = 0
    for k in range(1, 20):
        p, m, v, c = oadam.adam_update(p, m, v, g, c, lr=0.01)
        assert np.allclose(p, -0.01 * k * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)


def test_adam_matches_scalar_loop():
    r = np.random.default_rng(1)
    gs = r.standard_normal((30, 4))
    p = r.standard_normal(4)
    m = np.zeros(4)
    v = np.zeros(4)
    c = 0
    ps = p.copy()
    for t, g in enumerate(gs, start=1):
        p, m, v, c = oadam.adam_update(p, m, v, g, c, lr=3e-3, b1=0.8, b2=0.99, eps=1e-7)
    # independent scalar loop (Kingma & Ba, Algorithm 1)
    for j in range(4):
        mj = vj = 0.0
        pj = ps[j]
        for t, g in enumerate(gs[:, j], start=1):
            mj = 0.8 * mj + 0.2 * g
            vj = 0.99 * vj + 0.01 * g * g
            pj -= 3e-3 * (mj / (1 - 0.8 ** t)) / (math.sqrt(vj / (1 - 0.99 ** t)) + 1e-7)
        assert abs(pj - p[j]) < 1e-12


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


def test_ho_residual_vanishes_on_exact_solution():
    d, w0 = 2, 20
    req, C, L = opb.ho_soft(d, w0)
    t = torch.linspace(0.01, 0.99, 40, dtype=torch.float64)[:, None]
    u = torch.tensor(oex.ho_exact(t.numpy(), d, w0))
    h = 1e-5
    up = torch.tensor(oex.ho_exact(t.numpy() + h, d, w0))
    um = torch.tensor(oex.ho_exact(t.numpy() - h, d, w0))
    ut, utt = (up - um) / (2 * h), (up - 2 * u + um) / h ** 2
    xb = torch.zeros((1, 1), dtype=torch.float64)
    loss = L([[t, u, ut, utt], [xb, torch.ones_like(xb), torch.zeros_like(xb), torch.ones_like(xb),
                                torch.zeros_like(xb)]])
    assert loss.item() < 1e-4


def test_hard_bc_ansatz_values():
    x = torch.tensor([[0.0]], dtype=torch.float64)
    _, C, _ = opb.ho_hard(2, 80, 0.1)
    assert C(x, torch.tensor([[123.0]], dtype=torch.float64)).item() == 1.0
    _, Cb, _ = opb.burgers()
    xb = torch.tensor([[-1.0, 0.5], [1.0, 0.3], [0.3, 0.0]], dtype=torch.float64)
    ub = Cb(xb, torch.full((3, 1), 7.0, dtype=torch.float64))
    assert torch.allclose(ub[:2, 0], torch.zeros(2, dtype=torch.float64), atol=1e-15)
    assert abs(ub[2, 0].item() + math.sin(math.pi * 0.3)) < 1e-15
    _, Cp, _ = opb.poisson(0.1)
    xp = torch.tensor([[0.0, 0.4], [1.0, 0.2], [0.3, 0.0], [0.7, 1.0]], dtype=torch.float64)
    assert Cp(xp, torch.full((4, 1), 5.0, dtype=torch.float64)).abs().max() < 1e-15


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


def test_wave_loss_zero_for_plane_wave():
    """u = sin(x - c t)... satisfies u_xx + u_yy - u_tt / c^2 = 0."""
    _, _, L = opb.wave(np.array([[0.0, 0.0, 0.2, 1.0]]), c0=2.0)
    t = torch.linspace(0, 1, 10, dtype=torch.float64)[:, None]
    x = torch.cat([t, t, t], 1)
    uxx = -torch.sin(x[:, 0:1] - 2.0 * x[:, 2:3])
    utt = -4.0 * torch.sin(x[:, 0:1] - 2.0 * x[:, 2:3])
    assert L([[x, uxx, torch.zeros_like(uxx), utt]]).item() < 1e-30
