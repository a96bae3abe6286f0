"""Pins oracle/model.py: central finite differences of the model value
(first and second derivatives) and of the loss w.r.t. parameters (float64);
the POU invariant; the m = 1 (flag 0) reduction to a plain tanh MLP with a
hand-written derivative recursion; window closed forms; hand-derived
derivative trees of get_jmaps."""

import math

import numpy as np
import pytest
import torch

import setups
from oracle import model as om
from oracle import problems as opb
from oracle import takes as ot

D = torch.float64


def _value(ap, x):
    return setups.oracle_jets(ap, torch.as_tensor(x, dtype=torch.float32), x.shape[1], 0, 0)[0][0]


def _value64(ap, x64):
    """Model value at float64 points (takes from the float32-rounded points,
    which are identical for interior FD offsets used here)."""
    p = ap["static"]["decomposition"]["subdomain"]["params"]
    lo, hi = p[0].numpy().astype(np.float64), p[1].numpy().astype(np.float64)
    pou = ap["static"]["decomposition"]["subdomain"]["pou"].numpy().reshape(-1)
    takes, all_ims, *_ = ot.get_inputs(x64, np.ones(lo.shape[0], int), lo, hi, pou)
    layers = [(W.double()[all_ims], b.double()[all_ims]) for W, b in ap["trainable"]["network"]["subdomain"]["layers"]]
    dom = om.to_torch_dom(lo[all_ims], hi[all_ims], p[4].numpy()[all_ims], p[5].numpy()[all_ims], D)
    return om.fbpinn_model(layers, dom, torch.tensor(x64), takes, om.identity_constraint)[0][:, 0]


@pytest.mark.parametrize("nsub,lo,hi,ov,ls", [
    ((5,), (0.0,), (1.0,), 2.3, [1, 16, 1]),
    ((3, 3), (0.0, 0.0), (1.0, 1.0), 1.9, [2, 16, 16, 1]),
    ((2, 2, 2), (-1.0, -1.0, 0.0), (1.0, 1.0, 1.0), 1.9, [3, 16, 16, 1]),
])
def test_jets_match_central_differences(nsub, lo, hi, ov, ls):
    xd = len(lo)
    _, ap = setups.build(nsub, lo, hi, ov, ls, seed=1, unnorm=(0.2, 1.5))
    x = setups.uniform_points(40, [a + 0.05 for a in lo], [b - 0.05 for b in hi], seed=3)
    m1 = m2 = (1 << xd) - 1
    jets, takes, _ = setups.oracle_jets(ap, x, xd, m1, m2)
    x64 = x.double().numpy()
    h = 1e-4
    for i in range(xd):
        e = np.zeros(xd)
        e[i] = h
        fp, fm, f0 = _value64(ap, x64 + e), _value64(ap, x64 - e), _value64(ap, x64)
        d1 = (fp - fm) / (2 * h)
        d2 = (fp - 2 * f0 + fm) / (h * h)
        assert torch.allclose(jets[0], f0, rtol=0, atol=1e-12)
        s1 = jets[1 + i].abs().max()
        s2 = jets[1 + xd + i].abs().max()
        assert (jets[1 + i] - d1).abs().max() <= 1e-6 * s1, i
        assert (jets[1 + xd + i] - d2).abs().max() <= 1e-4 * s2, i


def test_pou_constant_invariant():
    """Every subdomain emitting the constant c gives u == c, zero derivatives."""
    _, ap = setups.build((4, 3), (0.0, 0.0), (1.0, 1.0), 1.9, [2, 8, 8, 1], seed=2)
    W, b = ap["trainable"]["network"]["subdomain"]["layers"][-1]
    ap["trainable"]["network"]["subdomain"]["layers"][-1] = (torch.zeros_like(W), torch.full_like(b, 0.37))
    x = setups.uniform_points(200, (0.0, 0.0), (1.0, 1.0), seed=5)
    jets, _, _ = setups.oracle_jets(ap, x, 2, 3, 3)
    c = float(torch.tensor(0.37, dtype=torch.float32))   # the float32 bias, exactly
    assert torch.allclose(jets[0], torch.full_like(jets[0], c), rtol=0, atol=1e-14)
    assert jets[1:].abs().max() < 1e-11


def _mlp_jets_by_hand(Ws, bs, x, mu, sd, umu, usd):
    """Independent forward-mode recursion (value, d/dx, d2/dx2) of
    u = (W_L tanh(... tanh(W1 (x-mu)/sd + b1) ...) + b_L) * usd + umu, 1-D."""
    a, da, dda = (x - mu) / sd, np.ones_like(x) / sd, np.zeros_like(x)
    a, da, dda = a[None], da[None], dda[None]
    for k, (W, b) in enumerate(zip(Ws, bs)):
        z = W @ a + b[:, None]
        dz = W @ da
        ddz = W @ dda
        if k < len(Ws) - 1:
            t = np.tanh(z)
            d1 = 1 - t * t
            a, da, dda = t, d1 * dz, d1 * ddz - 2 * t * d1 * dz * dz
        else:
            a, da, dda = z, dz, ddz
    return a[0] * usd + umu, da[0] * usd, dda[0] * usd


def test_single_subdomain_is_a_plain_mlp():
    _, ap = setups.build((1,), (0.0,), (1.0,), 1.0, [1, 12, 12, 1], seed=7, unnorm=(0.3, 2.0))
    p = ap["static"]["decomposition"]["subdomain"]["params"]
    assert p[4][0, 0] == 0  # flag 0: window switched off
    x = setups.uniform_points(64, (0.0,), (1.0,), seed=1)
    jets, _, _ = setups.oracle_jets(ap, x, 1, 1, 1)
    Ws = [W[0].double().numpy() for W, _ in ap["trainable"]["network"]["subdomain"]["layers"]]
    bs = [b[0].double().numpy() for _, b in ap["trainable"]["network"]["subdomain"]["layers"]]
    lo, hi = p[0][0, 0].double().item(), p[1][0, 0].double().item()
    umu, usd = float(p[5][0, 0]), float(p[5][0, 1])   # the stored float32 unnorm values
    u, du, ddu = _mlp_jets_by_hand(Ws, bs, x[:, 0].double().numpy(), (hi + lo) / 2, (hi - lo) / 2, umu, usd)
    assert np.allclose(jets[0].numpy(), u, rtol=1e-12, atol=1e-12)
    assert np.allclose(jets[1].numpy(), du, rtol=1e-11, atol=1e-11)
    assert np.allclose(jets[2].numpy(), ddu, rtol=1e-10, atol=1e-10)


def test_window_closed_forms():
    Ws = [torch.zeros(1, 1)]
    bs = [torch.ones(1)]
    xmin, xmax = torch.tensor([0.0, 1.0], dtype=D), torch.tensor([2.0, 3.0], dtype=D)
    one = torch.ones(1, dtype=D)
    for x, w_expect in ((torch.tensor([1.0, 2.0], dtype=D), 1.0), (torch.tensor([0.0, 2.0], dtype=D), 0.0),
                        (torch.tensor([0.5, 2.0], dtype=D), ((1 + math.cos(-math.pi / 2)) / 2) ** 2)):
        uw, w, _ = om._pair_fn([W.double() for W in [torch.zeros(1, 2)]], [torch.ones(1, dtype=D)], xmin, xmax,
                               one, torch.tensor([0.0, 1.0], dtype=D), x)
        assert abs(w.item() - w_expect) < 1e-15
        assert abs(uw.item() - w_expect) < 1e-15


def test_get_jmaps_hand_derived():
    nodes, leaves, jac_is = om.get_jmaps(((0, ()), (0, (0,)), (0, (0, 0))))
    assert nodes == (((0, 0), (0,), 0), ((1, 0), (0, 0), 1))
    assert leaves == ((2, (0, 0)),)
    assert jac_is == ((0, 0, 0), (0, 1, 0), (0, 2, 0))
    nodes, leaves, jac_is = om.get_jmaps(((0, (0, 0)), (0, (1, 1)), (0, (2, 2))))
    assert leaves == ((2, (0, 0)), (4, (1, 1)), (6, (2, 2)))
    assert jac_is == ((0, 2, 0), (1, 2, 0), (2, 2, 0))
    nodes, leaves, jac_is = om.get_jmaps(((0, ()),))
    assert nodes == () and leaves == ((0, ()),) and jac_is == ((0, 0, 0),)


def test_loss_gradient_central_differences():
    _, ap = setups.build((3,), (0.0,), (1.0,), 2.2, [1, 8, 8, 1], seed=4)
    x = setups.grid_points((30,), (0.0,), (1.0,))
    req, C, L = opb.ho_hard(2, 5, 0.1)
    takes, all_ims, layers, dom, _ = setups.oracle_view(ap, x)
    jm = [om.get_jmaps(r) for r in req]
    cons = [[x.double()]]
    loss, g, _ = om.value_and_grad(layers, dom, cons, [takes], jm, C, L)

    def lossat(ls):
        return om.fbpinn_loss(ls, dom, cons, [takes], jm, C, L).item()
    h = 1e-6
    r = np.random.default_rng(0)
    for li in range(len(layers)):
        for wb in range(2):
            t = layers[li][wb]
            for _ in range(3):
                idx = tuple(int(r.integers(0, s)) for s in t.shape)
                lp = [(W.clone(), b.clone()) for W, b in layers]
                lm = [(W.clone(), b.clone()) for W, b in layers]
                lp[li][wb][idx] += h
                lm[li][wb][idx] -= h
                fd = (lossat(lp) - lossat(lm)) / (2 * h)
                an = g[li][wb][idx].item()
                assert abs(fd - an) <= 1e-5 * max(1.0, abs(an)) + 1e-6 * abs(loss.item()) * 1e-3, (li, wb, idx, fd, an)
