"""Pins oracle/problems.py (the config problems' constraining_fn and loss_fn,
problems.py:68-481, and the multiscale Poisson problem).

Each loss_fn is evaluated on its problem's exact solution (oracle/exact.py,
pinned in tests/test_oracle_exact.py) with float64 central-difference
derivatives: it must vanish relative to the size of its terms, and the same
check must reject the exact fields with one derivative scaled by 1 + 1e-3, so
that a dropped term, a wrong sign or coefficient, or swapped derivative
columns fail it.  Loss terms with data (the inverse oscillator's 1e6
data weight, the soft boundary terms) and trainable coefficients are pinned
by their closed forms on shifted inputs.  Each hard-constraint ansatz is
evaluated through the oracle's nested jvp (the reference's derivative tree)
on an arbitrary smooth network output: the boundary / initial values and
derivatives it is built to enforce hold exactly, and its low-order Taylor
coefficients at t = 0 (which fix the widths sd and t1) match their hand
expansions.
"""

import math

import numpy as np
import torch

from oracle import exact as oex
from oracle import model as om
from oracle import problems as opb

D = torch.float64


def _nn(xd, seed=0, n=6):
    """An arbitrary smooth 'network output' R^xd -> R (float64)."""
    r = np.random.default_rng(seed)
    a = torch.tensor(r.standard_normal(n), dtype=D)
    b = torch.tensor(r.standard_normal((n, xd)) * 2.0, dtype=D)
    c = torch.tensor(r.uniform(0, 2 * math.pi, n), dtype=D)
    return lambda x: (torch.sin(x @ b.T + c) * a).sum(1, keepdim=True) + 0.3


def _jets(constraining_fn, nn, x, req):
    """Jets of constraining_fn(x, nn(x)) through the oracle's nested jvp."""
    jm = om.get_jmaps(tuple(req))
    return om.get_ujs(x, jm, lambda xb: (constraining_fn(xb, nn(xb)), ()))


def _rel(loss, scale):
    return float(loss) / float(scale)


# ---------------------------------------------------------------- oscillators

def _ho_fields(d, w0, t, h=1e-5):
    f = lambda s: oex.ho_exact(s, d, w0)
    u = f(t)
    ut = (f(t + h) - f(t - h)) / (2 * h)
    utt = (f(t + h) - 2 * u + f(t - h)) / h ** 2
    T = lambda a: torch.tensor(a, dtype=D)[:, None]
    return T(t), T(u), T(ut), T(utt)


def test_ho_hard_loss_vanishes_on_exact_solution():
    d, w0 = 2, 80
    _, _, L = opb.ho_hard(d, w0, 0.1)
    t, u, ut, utt = _ho_fields(d, w0, np.linspace(0.01, 0.99, 200))
    scale = torch.mean((w0 ** 2 * u) ** 2)
    assert _rel(L([[t, u, ut, utt]]), scale) < 1e-12
    for bad in ([t, u * 1.001, ut, utt], [t, u, ut * 1.001, utt], [t, u, ut, utt * 1.001]):
        assert _rel(L([bad]), scale) > 1e-10


def test_ho_hard_ansatz_enforces_initial_conditions():
    sd = 0.1
    _, C, _ = opb.ho_hard(2, 80, sd)
    nn = _nn(1)
    x = torch.tensor([[0.0], [1.0]], dtype=D)
    u, ut, utt = _jets(C, nn, x, [(0, ()), (0, (0,)), (0, (0, 0))])
    assert u[0, 0].item() == 1.0 and ut[0, 0].item() == 0.0          # u(0) = 1, u'(0) = 0
    # tanh(t/sd)^2 = t^2/sd^2 + O(t^4): u''(0) = 2 NN(0) / sd^2
    assert abs(utt[0, 0].item() - 2 * nn(x[:1])[0, 0].item() / sd ** 2) < 1e-9 * abs(utt[0, 0].item())
    # far from t = 0 the ansatz is 1 + NN
    assert abs(u[1, 0].item() - 1 - nn(x[1:])[0, 0].item()) < 1e-7


def test_ho_soft_loss_vanishes_on_exact_solution_and_weights_boundary():
    d, w0 = 2, 20
    _, _, L = opb.ho_soft(d, w0)
    t, u, ut, utt = _ho_fields(d, w0, np.linspace(0.01, 0.99, 100))
    xb = torch.zeros((1, 1), dtype=D)
    one, zero = torch.ones_like(xb), torch.zeros_like(xb)
    scale = torch.mean((w0 ** 2 * u) ** 2)
    assert _rel(L([[t, u, ut, utt], [xb, one, zero, one, zero]]), scale) < 1e-12
    assert _rel(L([[t, u, ut * 1.001, utt], [xb, one, zero, one, zero]]), scale) > 1e-10
    # boundary terms: 1e6 (u - 1)^2 + 1e2 (u_t - 0)^2
    phys = L([[t, u, ut, utt], [xb, one, zero, one, zero]]).item()
    eps, eta = 1e-3, 0.2
    val = L([[t, u, ut, utt], [xb, one, zero, one + eps, zero + eta]]).item() - phys
    assert abs(val - (1e6 * eps ** 2 + 1e2 * eta ** 2)) < 1e-9


def test_ho_inverse_loss_closed_forms():
    d, w0 = 2, 80
    _, _, L = opb.ho_inverse(w0 ** 2)
    t, u, ut, utt = _ho_fields(d, w0, np.linspace(0.01, 0.99, 200))
    xo = torch.linspace(0, 1, 13, dtype=D)[:, None]
    uo = torch.tensor(oex.ho_exact(xo.numpy(), d, w0), dtype=D)
    scale = torch.mean((w0 ** 2 * u) ** 2)
    mu_true = torch.tensor(2.0 * d, dtype=D)
    assert _rel(L([[t, u, ut, utt], [xo, uo, uo]], mu_true), scale) < 1e-12
    # the trainable coefficient multiplies u_t: loss(mu_true + delta) = delta^2 mean(u_t^2) + O(FD error)
    delta = 0.5
    got = L([[t, u, ut, utt], [xo, uo, uo]], mu_true + delta).item()
    want = delta ** 2 * torch.mean(ut ** 2).item()
    assert abs(got - want) < 1e-4 * want
    # and its gradient is 2 delta mean(u_t^2)
    mu = (mu_true + delta).clone().requires_grad_(True)
    g, = torch.autograd.grad(L([[t, u, ut, utt], [xo, uo, uo]], mu), [mu])
    assert abs(g.item() - 2 * delta * torch.mean(ut ** 2).item()) < 1e-4 * abs(g.item())
    # data term: 1e6 mean((u - u_obs)^2)
    eps = torch.full_like(uo, 1e-3)
    eps[::2] = -2e-3
    got = L([[t, u, ut, utt], [xo, uo + eps, uo]], mu_true).item()
    want = 1e6 * torch.mean(eps ** 2).item()
    assert abs(got - want) < 1e-6 * want


# ---------------------------------------------------------------- Burgers

def test_burgers_loss_vanishes_on_exact_solution():
    nu = 0.01 / math.pi
    _, _, L = opb.burgers(nu, 0.1)
    xs = np.array([-0.9, -0.75, -0.6, -0.45, 0.45, 0.6, 0.75, 0.9])
    ts = np.array([0.1, 0.2, 0.3, 0.4])
    h = k = 1e-4
    uf = lambda xx, tt: oex.burgers_exact(nu, xx, [tt])[:, 0]
    rows = []
    for t0 in ts:
        u0 = uf(xs, t0)
        ux = (uf(xs + h, t0) - uf(xs - h, t0)) / (2 * h)
        uxx = (uf(xs + h, t0) - 2 * u0 + uf(xs - h, t0)) / h ** 2
        ut = (uf(xs, t0 + k) - uf(xs, t0 - k)) / (2 * k)
        rows.append(np.stack([xs, np.full_like(xs, t0), u0, ux, ut, uxx], 1))
    a = torch.tensor(np.concatenate(rows), dtype=D)
    x, u, ux, ut, uxx = a[:, 0:2], a[:, 2:3], a[:, 3:4], a[:, 4:5], a[:, 5:6]
    scale = torch.mean(ut ** 2)
    assert _rel(L([[x, u, ux, ut, uxx]]), scale) < 1e-9
    for bad in ([x, u * 1.001, ux, ut, uxx], [x, u, ux * 1.001, ut, uxx], [x, u, ux, ut * 1.001, uxx],
                [x, u, ut, ux, uxx]):
        assert _rel(L([bad]), scale) > 1e-8
    # the viscosity enters as nu = 0.01/pi: the same fields fail with 0.01
    _, _, L2 = opb.burgers(0.01, 0.1)
    assert _rel(L2([[x, u, ux, ut, uxx]]), scale) > 1e-6


def test_burgers_ansatz_enforces_initial_and_boundary_values():
    _, C, _ = opb.burgers(0.01 / math.pi, 0.1)
    nn = _nn(2, seed=1)
    xs = torch.linspace(-0.9, 0.9, 7, dtype=D)
    x0 = torch.stack([xs, torch.zeros_like(xs)], 1)
    u, ux, uxx = _jets(C, nn, x0, [(0, ()), (0, (0,)), (0, (0, 0))])
    s = torch.sin(math.pi * xs)[:, None]
    c = torch.cos(math.pi * xs)[:, None]
    assert (u + s).abs().max() < 1e-15                       # u(x, 0) = -sin(pi x)
    assert (ux + math.pi * c).abs().max() < 1e-12
    assert (uxx - math.pi ** 2 * s).abs().max() < 1e-10
    tb = torch.linspace(0.1, 1.0, 5, dtype=D)
    xb = torch.cat([torch.stack([torch.full_like(tb, -1.0), tb], 1), torch.stack([torch.full_like(tb, 1.0), tb], 1)])
    u, ut = _jets(C, nn, xb, [(0, ()), (0, (1,))])
    assert u.abs().max() < 1e-15 and ut.abs().max() < 1e-15   # u(+-1, t) = 0 for all t


# ---------------------------------------------------------------- Poisson

def test_poisson_source_hand_values():
    x = torch.tensor([[0.25, 0.25], [0.125, 0.125], [1 / 16, 1 / 16], [0.5, 0.5], [0.25, 0.75]], dtype=D)
    f = opb.poisson_source(x)[:, 0]
    want = [1.0, 1.5, math.sin(math.pi / 8) ** 2 + 0.5 + 1.0, 0.0, -1.0]
    assert torch.allclose(f, torch.tensor(want, dtype=D), rtol=0, atol=1e-14)


def test_poisson_loss_vanishes_on_exact_solution():
    _, _, L = opb.poisson(0.1)
    r = np.random.default_rng(3)
    xn = r.uniform(0.02, 0.98, (300, 2))
    h = 1e-4
    d2 = []
    for i in range(2):
        e = np.zeros(2)
        e[i] = h
        d2.append((oex.poisson_exact(xn + e) - 2 * oex.poisson_exact(xn) + oex.poisson_exact(xn - e)) / h ** 2)
    x = torch.tensor(xn, dtype=D)
    uxx, uyy = torch.tensor(d2[0], dtype=D), torch.tensor(d2[1], dtype=D)
    scale = torch.mean(opb.poisson_source(x) ** 2)
    assert _rel(L([[x, uxx, uyy]]), scale) < 1e-9
    assert _rel(L([[x, uxx * 1.001, uyy]]), scale) > 1e-8
    assert _rel(L([[x, -uxx, -uyy]]), scale) > 1.0


def test_poisson_ansatz_vanishes_on_the_boundary():
    _, C, _ = opb.poisson(0.1)
    nn = _nn(2, seed=2)
    s = torch.linspace(0.05, 0.95, 9, dtype=D)
    z, o = torch.zeros_like(s), torch.ones_like(s)
    for pts, tang in ((torch.stack([z, s], 1), 1), (torch.stack([o, s], 1), 1),
                      (torch.stack([s, z], 1), 0), (torch.stack([s, o], 1), 0)):
        u, ut = _jets(C, nn, pts, [(0, ()), (0, (tang,))])
        assert u.abs().max() < 1e-15 and ut.abs().max() < 1e-15
    # inside, five widths sd from every edge, the ansatz is the network output to 1e-3
    xc = torch.tensor([[0.5, 0.5]], dtype=D)
    assert abs(C(xc, nn(xc)).item() - nn(xc).item()) < 1e-3 * abs(nn(xc).item())


# ---------------------------------------------------------------- wave

SOURCES = np.array([[0.1, -0.2, 0.3, 1.0], [-0.3, 0.25, 0.1, 0.5], [0.2, 0.3, 0.03, -0.8]])


def test_wave_ansatz_initial_condition():
    """u(x, y, 0) = sum of the Gaussian sources and u_t(x, y, 0) = 0 for any
    network output; u_tt(x, y, 0) = (-2.25 f + 12.5 NN) / t1^2 with
    t1 = min sigma / c0 (Taylor coefficients of exp(-(1.5 t/t1)^2 / 2) and
    tanh(2.5 t/t1)^2); u = NN once t >> t1."""
    c0 = 1.3
    _, C, _ = opb.wave(SOURCES, c0=c0)
    nn = _nn(3, seed=4)
    r = np.random.default_rng(5)
    xy = r.uniform(-1, 1, (40, 2))
    x = torch.tensor(np.concatenate([xy, np.zeros((40, 1))], 1), dtype=D)
    u, ut, utt = _jets(C, nn, x, [(0, ()), (0, (2,)), (0, (2, 2))])
    f = np.zeros(40)
    for k in range(40):                                   # independent per-point loop
        for cx, cy, s, a in SOURCES:
            f[k] += a * math.exp(-0.5 * ((xy[k, 0] - cx) ** 2 + (xy[k, 1] - cy) ** 2) / s ** 2)
    assert np.abs(u[:, 0].numpy() - f).max() < 1e-14
    assert ut.abs().max() < 1e-13
    t1 = SOURCES[:, 2].min() / c0
    want = (-2.25 * f + 12.5 * nn(x)[:, 0].numpy()) / t1 ** 2
    assert np.abs(utt[:, 0].numpy() - want).max() < 1e-9 * np.abs(want).max()
    xl = x.clone()
    xl[:, 2] = 1.0
    assert (C(xl, nn(xl)) - nn(xl)).abs().max() < 1e-12


def test_wave_loss_zero_for_plane_wave():
    """u = sin(x - c t) satisfies u_xx + u_yy - u_tt / c^2 = 0; with c = 2 a
    missing 1/c^2 would leave 3 u_xx."""
    _, _, L = opb.wave(np.array([[0.0, 0.0, 0.2, 1.0]]), c0=2.0)
    t = torch.linspace(0, 1, 10, dtype=D)[:, None]
    x = torch.cat([t, t, t], 1)
    uxx = -torch.sin(x[:, 0:1] - 2.0 * x[:, 2:3])
    utt = -4.0 * torch.sin(x[:, 0:1] - 2.0 * x[:, 2:3])
    assert L([[x, uxx, torch.zeros_like(uxx), utt]]).item() < 1e-30
    assert L([[x, uxx, torch.zeros_like(uxx), utt * 1.001]]).item() > 1e-9


def test_wave_loss_vanishes_on_free_space_solution():
    c0 = 1.0
    src = SOURCES[:2]
    _, _, L = opb.wave(src, c0=c0)
    r = np.random.default_rng(6)
    n = 150
    xy = r.uniform(-0.9, 0.9, (n, 2))
    t = r.uniform(0.05, 0.9, n)
    h = 1e-3
    f = lambda dx, dy, dt: oex.wave_free_space(xy + [dx, dy], t + dt, src, c0=c0)
    u0 = f(0, 0, 0)
    uxx = (f(h, 0, 0) - 2 * u0 + f(-h, 0, 0)) / h ** 2
    uyy = (f(0, h, 0) - 2 * u0 + f(0, -h, 0)) / h ** 2
    utt = (f(0, 0, h) - 2 * u0 + f(0, 0, -h)) / h ** 2
    T = lambda a: torch.tensor(a, dtype=D)[:, None]
    x = torch.tensor(np.concatenate([xy, t[:, None]], 1), dtype=D)
    scale = torch.mean(T(utt) ** 2)
    assert _rel(L([[x, T(uxx), T(uyy), T(utt)]]), scale) < 1e-8
    assert _rel(L([[x, T(uxx), T(uyy) * 1.001, T(utt)]]), scale) > 1e-8
