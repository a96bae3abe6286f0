"""Full-size parity: the (2+1)D wave (12 x 12 x 6 subdomains, 1,048,576
points) and multiscale Poisson (16 x 16, 262,144 points) configurations run
with their production launch shape (autotuned persistent bins or chunks, as
bench.py runs them), checked against the float64 oracle on parts of the
problem small enough for it:

  * jets: the per-point jets of the full forward (FBPINNJets over the whole
    work list) at 1,500 random points against the oracle's nested jvp at
    those points (a point's jets depend only on the subdomains holding it),
    1e-5 relative per jet column;
  * gradient rows: d loss / d theta_k of two random subdomains k from the full
    backward against the oracle's gradient of the loss over the points inside
    box k only, scaled by (points in box k) / (all points): the points
    outside box k do not depend on theta_k (trainers.py:126-177, 249-267),
    1e-4 relative per layer tensor.
"""

import numpy as np
import pytest
import torch

from fbpinns_amd import configs
from fbpinns_amd.ops import FBPINNJets
from fbpinns_amd.trainers import FBPINNTrainer

import setups
from oracle import model as om
from oracle import problems as opb
from oracle import takes as ot

pytestmark = pytest.mark.gpu


def _oracle_problem(name, ap):
    prob = ap["static"]["problem"]
    if name == "wave":
        return opb.wave(prob["source"].cpu().numpy(), prob["c0"])
    return opb.poisson(prob["sd"], prob["ks"])


def _box_gradient_row(eng, name, k):
    """Oracle d loss / d theta_k from the points inside box k, as the layer
    tensors of row k, and the number of those points."""
    ap = eng.all_params
    p = ap["static"]["decomposition"]["subdomain"]["params"]
    xmins, xmaxs = p[0].cpu().numpy(), p[1].cpu().numpy()
    pou = ap["static"]["decomposition"]["subdomain"]["pou"].cpu().numpy().reshape(-1)
    x = eng.x.cpu().numpy()
    inside = np.all((x >= xmins[k]) & (x <= xmaxs[k]), axis=1)
    xk = x[inside]
    req, C, L = _oracle_problem(name, ap)
    takes, all_ims, *_ = ot.get_inputs(xk, np.ones(eng.m, dtype=int), xmins, xmaxs, pou)
    layers = [(W.detach().cpu()[all_ims].double(), b.detach().cpu()[all_ims].double())
              for W, b in ap["trainable"]["network"]["subdomain"]["layers"]]
    dom = om.to_torch_dom(xmins[all_ims], xmaxs[all_ims], p[4].cpu().numpy()[all_ims],
                          p[5].cpu().numpy()[all_ims], torch.float64)
    cons = [[torch.tensor(xk, dtype=torch.float64)]]
    _, gl, _ = om.value_and_grad(layers, dom, cons, [takes], [om.get_jmaps(r) for r in req], C, L)
    row = int(np.flatnonzero(all_ims == k)[0])
    scale = xk.shape[0] / x.shape[0]
    return [(gW[row] * scale, gb[row] * scale) for gW, gb in gl], xk.shape[0]


@pytest.mark.parametrize("name", ["poisson", "wave"])
def test_full_size_production_launch_matches_oracle(name):
    c = configs.CONFIGS[name]()
    tr = FBPINNTrainer(c, write_outputs=False)
    eng, sched, _, _ = tr.setup(exact=False)
    eng.set_active(next(iter(sched)))
    wl = eng.plan.wl
    print(f"[full] {name}: {wl.n_points} points, {wl.n_pairs} pairs, {wl.n_chunks} chunks, launch {eng.launch}, "
          f"workgroups {wl.n_fbins}/{wl.n_bbins}", flush=True)
    assert wl.n_points == int(np.prod(c.ns[0]))

    # jets of the full forward at random points
    with torch.no_grad():
        ujet = FBPINNJets.apply(eng.theta, eng.plan)
    torch.cuda.synchronize()
    r = np.random.default_rng(7)
    sel = np.sort(r.choice(wl.n_points, 1500, replace=False))
    xs = eng.x[torch.as_tensor(sel, device=eng.device)].cpu()
    ref, _, _ = setups.oracle_jets(eng.all_params, xs, eng.net.xd, eng.net.d1mask, eng.net.d2mask)
    err = setups.col_rel_err(ujet[:, torch.as_tensor(sel, device=eng.device)], ref)
    print(f"  jets at {len(sel)} points: per-column rel err {[f'{e:.1e}' for e in err.tolist()]}", flush=True)
    assert err.max().item() <= 1e-5

    # gradient rows of two random subdomains from the full backward
    loss, g, _ = eng.loss_and_grad()
    torch.cuda.synchronize()
    g = g.detach().cpu()
    for k in r.choice(eng.m, 2, replace=False):
        orow, nk = _box_gradient_row(eng, name, int(k))
        off = 0
        errs = []
        for gW, gb in orow:
            for o in (gW, gb):
                n = o.numel()
                kr = g[k, off:off + n].reshape(o.shape).double()
                off += n
                errs.append(((kr - o).abs().max() / o.abs().max().clamp_min(1e-30)).item())
        print(f"  subdomain {int(k)}: {nk} points in its box, grad rel err per tensor "
              f"{[f'{e:.1e}' for e in errs]}", flush=True)
        assert max(errs) <= 1e-4
