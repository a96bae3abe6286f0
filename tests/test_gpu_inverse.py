"""GPU: many independent inverse problems in one launch (config 5 style).

Gradients of the batched loss w.r.t. subdomain parameters and every
instance's learnable mu are checked against (a) float64 autograd through the
oracle and (b) float64 central finite differences of the oracle loss.
Tolerance: max |gpu - oracle64| <= max(1e-4, 3 e32) * max |oracle64| per
tensor, where e32 is the same error of the oracle run in float32 on the same
inputs: with w0 up to 100 the residual u_tt + mu u_t + k u cancels terms of
size ~1e4, so float32 itself carries ~1e-4 relative error in d loss / d mu.
Instances do not interact: perturbing one instance's mu changes only its own
gradient entries.
"""

import numpy as np
import pytest
import torch

from fbpinns_amd import configs
from fbpinns_amd.trainers import FBPINNTrainer

from oracle import model as om
from oracle import takes as ot

pytestmark = pytest.mark.gpu


def _engine(n_inst, graph=False, launch=None):
    c = configs.inverse_batch(n_inst=n_inst, graph=graph)
    tr = FBPINNTrainer(c, write_outputs=False)
    eng, sched, x_test, u_exact = tr.setup()
    eng.set_active(next(iter(sched)), launch=launch)
    return c, eng, x_test, u_exact


def _oracle(eng):
    ap = eng.all_params
    p = ap["static"]["decomposition"]["subdomain"]["params"]
    lo, hi = p[0].cpu().numpy(), p[1].cpu().numpy()
    pou = ap["static"]["decomposition"]["subdomain"]["pou"].cpu().numpy().reshape(-1)
    m = lo.shape[0]
    cons_g = [[v.cpu() for v in c] for c in eng.cglobal]
    xg = torch.cat([c[0] for c in cons_g]).numpy()
    sizes = [c[0].shape[0] for c in cons_g]
    offs = list(np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(int))
    tips, local, cfs, cips = ot.get_x_batch(np.ones(m, int), xg, offs, sizes, lo, hi)
    takes, all_ims, *_ = ot.get_inputs(xg[tips], np.ones(m, int), lo, hi, pou)
    takess = ot.split_takes(takes, cfs, cips, len(tips))
    cons = []
    for c, loc in zip(cons_g, local):
        cons.append([v[torch.as_tensor(loc)].double() if v.is_floating_point() else v[torch.as_tensor(loc)]
                     for v in c])
    layers = [(W.detach().cpu()[all_ims].double(), b.detach().cpu()[all_ims].double())
              for W, b in ap["trainable"]["network"]["subdomain"]["layers"]]
    dom = om.to_torch_dom(lo[all_ims], hi[all_ims], p[4].cpu().numpy()[all_ims], p[5].cpu().numpy()[all_ims],
                          torch.float64)
    prob = ap["static"]["problem"]
    k = prob["k"].detach().cpu().double()
    n = prob["n_inst"]

    def loss_fn(cs, mu):
        _, inst, u, ut, utt = cs[0]
        r = utt + mu[inst][:, None] * ut + k[inst][:, None] * u
        _, uc, _, u_o = cs[1]
        return n * torch.mean(r ** 2) + 1e6 * n * torch.mean((u_o - uc) ** 2)

    jm = [om.get_jmaps(r) for r in eng.required_ujss]
    return layers, dom, cons, takess, jm, loss_fn, all_ims


def test_batched_gradients_vs_float64_oracle_and_fd():
    c, eng, _, _ = _engine(n_inst=6)
    # move mu off zero so its gradient is generic
    eng.pflat.copy_(torch.linspace(0.5, 3.0, eng.Q, device=eng.device))
    print(f"[inverse] m={eng.m} pairs={eng.plan.wl.n_pairs} Q={eng.Q}", flush=True)
    loss, g, gp = eng.loss_and_grad()
    torch.cuda.synchronize()
    layers, dom, cons, takess, jm, loss_fn, all_ims = _oracle(eng)
    mu64 = eng.pflat.detach().cpu().double()[:eng.Q]
    oloss, og, ogx = om.value_and_grad(layers, dom, cons, takess, jm, om.identity_constraint, loss_fn,
                                       extra=(mu64,))
    # the oracle's own float32 error on the same inputs (conditioning yardstick)
    f32 = lambda t: t.float() if t.is_floating_point() else t
    l32 = [(f32(W), f32(b)) for W, b in layers]
    d32 = {k_: f32(v_) for k_, v_ in dom.items()}
    c32 = [[f32(v_) for v_ in c_] for c_ in cons]
    _, og32, ogx32 = om.value_and_grad(l32, d32, c32, takess, jm, om.identity_constraint, loss_fn,
                                       extra=(mu64.float(),))
    e32 = max([((a.double() - b).abs().max() / b.abs().max()).item() for (a, _), (b, _) in zip(og32, og)]
              + [((ogx32[0].double() - ogx[0]).abs().max() / ogx[0].abs().max()).item()])
    tol = max(1e-4, 3 * e32)
    print(f"  oracle float32 error {e32:.2e} -> tolerance {tol:.2e}")
    el = abs(loss.item() - oloss.item()) / abs(oloss.item())
    print(f"  loss gpu {loss.item():.8e} oracle {oloss.item():.8e} rel {el:.2e}")
    assert el <= 1e-4
    g = g.detach().cpu()
    off = 0
    for (gW, gb) in og:
        o, i = gW.shape[1], gW.shape[2]
        for ref, sl in ((gW, (slice(off, off + o * i), (o, i))), (gb, (slice(off + o * i, off + o * i + o), (o,)))):
            a = g[all_ims][:, sl[0]].reshape((len(all_ims),) + sl[1]).double()
            e = ((a - ref).abs().max() / ref.abs().max()).item()
            print(f"  d{tuple(ref.shape)} rel {e:.2e}")
            assert e <= tol
        off += o * i + o
    emu = ((gp[:eng.Q].cpu().double() - ogx[0]).abs().max() / ogx[0].abs().max()).item()
    print(f"  dmu rel {emu:.2e}")
    assert emu <= tol

    # float64 central differences of the oracle loss
    def L(ls, mu):
        return om.fbpinn_loss(ls, dom, cons, takess, jm, om.identity_constraint,
                              lambda cs: loss_fn(cs, mu)).item()
    # h balances truncation (h^2) against rounding (eps * loss / h, loss ~ 3e7)
    h = 1e-4
    for j in range(eng.Q):
        mp, mm = mu64.clone(), mu64.clone()
        mp[j] += h
        mm[j] -= h
        fd = (L(layers, mp) - L(layers, mm)) / (2 * h)
        assert abs(fd - ogx[0][j].item()) <= 1e-5 * abs(ogx[0]).max().item(), (j, fd, ogx[0][j].item())
    r = np.random.default_rng(0)
    for li, (W, b) in enumerate(layers):
        for _ in range(4):
            idx = tuple(int(r.integers(0, s)) for s in W.shape)
            lp = [(A.clone(), B.clone()) for A, B in layers]
            lm = [(A.clone(), B.clone()) for A, B in layers]
            lp[li][0][idx] += h
            lm[li][0][idx] -= h
            fd = (L(lp, mu64) - L(lm, mu64)) / (2 * h)
            row = all_ims[idx[0]]
            off = sum(a.shape[1] * a.shape[2] + a.shape[1] for a, _ in layers[:li])
            kern = g[row, off + idx[1] * W.shape[2] + idx[2]].item()
            scale = og[li][0].abs().max().item()
            assert abs(fd - og[li][0][idx].item()) <= 1e-5 * scale, (li, idx, fd)
            assert abs(fd - kern) <= tol * scale, (li, idx, fd, kern)


def test_instances_are_independent():
    c, eng, _, _ = _engine(n_inst=6)
    _, g0, gp0 = eng.loss_and_grad()
    eng.pflat[2] += 0.5           # perturb instance 2's mu only
    _, g1, gp1 = eng.loss_and_grad()
    changed = (gp1 - gp0).abs() > 0
    assert changed[2] and changed.sum().item() == 1
    m_per = [len(x[0]) for x in c.decomposition_init_kwargs["subdomain_xss"]]
    lo = sum(m_per[:2])
    rows = (g1 - g0).abs().sum(1) > 0
    assert rows[lo:lo + m_per[2]].all() and rows.sum().item() == m_per[2]


def test_256_instances_wave_bins_match_workgroup_bins():
    """The 256-instance work list (some 6,000 chunks of about 23 pairs)
    runs one wave per chunk by default; against one workgroup per 512-pair
    chunk its loss agrees bit for bit and its gradients (network rows and the
    256 learnable mu) to float32 summation accuracy."""
    _, ew, _, _ = _engine(n_inst=256)
    _, ec, _, _ = _engine(n_inst=256, launch=(("chunks", 512), ("chunks", 512)))
    print(f"[inverse256 launch] default {ew.launch} bins {ew.plan.wl.n_fbins}, workgroups "
          f"{ew.plan.wl.workgroups()}; chunks {ec.plan.wl.n_chunks}", flush=True)
    assert ew.launch == (("waves", 32), ("waves", 32))
    lw, gw, pw = ew.loss_and_grad()
    lc, gc, pc = ec.loss_and_grad()
    e = ((gw - gc).abs().max() / gc.abs().max()).item()
    ep = ((pw - pc).abs().max() / pc.abs().max()).item()
    print(f"  loss {lw.item():.8e} / {lc.item():.8e}; grad rel {e:.2e}, mu grad rel {ep:.2e}", flush=True)
    assert lw.item() == lc.item()
    assert e < 1e-5 and ep < 1e-5


def test_256_instances_train_in_one_launch():
    c, eng, x_test, u_exact = _engine(n_inst=256, graph=True)
    mu_true = torch.tensor(c.problem_init_kwargs["mus_true"], dtype=torch.float32)
    print(f"[inverse256] m={eng.m} points={eng.plan.wl.n_points} pairs={eng.plan.wl.n_pairs} "
          f"chunks={eng.plan.wl.n_chunks}", flush=True)
    l0 = eng.step().item()
    err0 = (eng.pflat[:256].cpu() - mu_true).abs().mean().item()
    for _ in range(3000):
        loss = eng.step()
    l1 = loss.item()
    err1 = (eng.pflat[:256].cpu() - mu_true).abs().mean().item()
    print(f"  loss {l0:.4e} -> {l1:.4e}; mean |mu - mu_true| {err0:.3f} -> {err1:.3f}", flush=True)
    assert l1 < l0 and err1 < err0
