"""GPU: the affine constraining-function kernel (fbpx_affine_jets and its
transpose) against the nested-jvp derivatives of constraining_fn evaluated
in float64 on the CPU.

Forward: outputs within 1e-5 of the float64 reference relative to each
output's max magnitude.  Backward: the gradient w.r.t. the u~ jets equals
float64 autograd through the generic (nested jvp) path, same tolerance.
Two constraints share one jet array (disjoint point ranges).
"""

import pytest
import torch

from fbpinns_amd import epilogue
from fbpinns_amd import problems as P
from fbpinns_amd.engine import pick_schema
from fbpinns_amd.ops import AffineJets, NetSpec

pytestmark = pytest.mark.gpu

CASES = [
    (P.HarmonicOscillator1DHardBC, dict(d=2, w0=20), 1,
     [((0, ()), (0, (0,)), (0, (0, 0))), ((0, ()),)]),
    (P.BurgersEquation2D, dict(), 2, [((0, ()), (0, (0,)), (0, (1,)), (0, (0, 0))), ((0, (1,)),)]),
    (P.WaveEquationConstantVelocity3D, dict(c0=1.3, source=P.wave_sources()), 3,
     [((0, (0, 0)), (0, (1, 1)), (0, (2, 2))), ((0, ()), (0, (2,)))]),
    (P.PoissonMultiscale2D, dict(), 2, [((0, (0, 0)), (0, (1, 1))), ((0, ()),)]),
]


def _params(prob, kw):
    s, t = prob.init_params(**kw)
    return {"static": {"problem": s}, "trainable": {"problem": t}}


@pytest.mark.parametrize("prob,kw,xd,reqs", CASES, ids=["ho", "burgers", "wave", "poisson"])
def test_affine_jets_match_nested_jvp(prob, kw, xd, reqs):
    torch.manual_seed(0)
    sizes = [300, 77]
    n = sum(sizes)
    cranges = [(0, sizes[0]), (sizes[0], sizes[1])]
    x64 = torch.rand((n, xd), dtype=torch.float64) * 0.8 + 0.1
    d1, d2 = pick_schema(xd, *epilogue.schema_masks(reqs, xd))
    net = NetSpec(xd, 32, 2, d1, d2)
    ap = _params(prob, kw)
    dev = torch.device("cuda")
    ap_gpu = {"static": {"problem": {k: (v.to(dev) if isinstance(v, torch.Tensor) else v)
                                     for k, v in ap["static"]["problem"].items()}},
              "trainable": ap["trainable"]}
    x32 = x64.float().to(dev)
    ep = epilogue.Epilogue(prob, ap_gpu, x32, cranges, [[x32[o:o + s]] for o, s in cranges], reqs, net)
    assert ep.affine is not None
    q64 = torch.randn((net.J, n), dtype=torch.float64)
    q = q64.float().to(dev).requires_grad_(True)
    A, B = ep.affine
    codes = [ep._req_codes(r) for r in ep.required_ujss]
    print(f"[affine] {prob.__name__} J={net.J} codes={codes}", flush=True)
    outs = AffineJets.apply(q, A, B, net.as_list(), cranges, codes)

    # float64 reference: nested jvp of constraining_fn on the Taylor stand-in
    ep64 = epilogue.Epilogue(prob, ap, x64, cranges, [[x64[o:o + s]] for o, s in cranges], reqs, net)
    q64r = q64.clone().requires_grad_(True)
    ref = []
    for ic, (off, n_c) in enumerate(cranges):
        ref += ep64._ujs_generic(ic, q64r[:, off:off + n_c], ap)
    assert len(ref) == len(outs)
    for o, r in zip(outs, ref):
        e = ((o.detach().cpu().double() - r.detach()).abs().max() / r.detach().abs().max().clamp_min(1e-30)).item()
        assert e < 1e-5, e

    # backward: random output gradients
    gouts = [torch.randn(o.shape, dtype=torch.float64) for o in outs]
    (gq,) = torch.autograd.grad(outs, (q,), [g.float().to(dev) for g in gouts])
    (gq64,) = torch.autograd.grad(ref, (q64r,), gouts)
    e = ((gq.cpu().double() - gq64).abs().max() / gq64.abs().max()).item()
    print(f"  backward rel err {e:.2e}", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("prob,kw,xd,lo,hi,nsub,ls", [
    (P.WaveEquationConstantVelocity3D, dict(c0=1.0, source=P.wave_sources()), 3, (-1.0, -1.0, 0.0),
     (1.0, 1.0, 1.0), (3, 3, 2), [3, 64, 64, 1]),
    (P.BurgersEquation2D, dict(), 2, (-1.0, 0.0), (1.0, 1.0), (4, 3), [2, 32, 32, 1]),
], ids=["wave", "burgers"])
def test_pou_reduce_with_affine_map_equals_separate_kernels(prob, kw, xd, lo, hi, nsub, ls):
    """fbpx_pou_reduce with the affine jets A, B equals fbpx_pou_reduce
    followed by fbpx_affine_jets on every column (bit for bit: same
    operations), and fbpx_pou_backward with B equals fbpx_affine_jets_backward
    followed by fbpx_pou_backward; its g_scale equals scaling the incoming
    jets first, bit for bit."""
    import numpy as np
    import setups
    from fbpinns_amd import ops
    from fbpinns_amd.engine import FusedEngine
    from fbpinns_amd.networks import FCN
    dec_cls, ap = setups.build(nsub, lo, hi, 1.9, ls, seed=0)
    ap["static"]["problem"] = prob.init_params(**kw)[0]
    eng = FusedEngine(prob, dec_cls, FCN, ap, device="cuda:0", graph=False)
    x = setups.uniform_points(3000, lo, hi, seed=4).cuda()
    full = (1 << xd) - 1
    plan, _ = eng.make_plan(x, schema=pick_schema(xd, full, full))
    net = plan.net
    print(f"[affine pou] {prob.__name__} J={net.J} pairs={plan.wl.n_pairs}", flush=True)
    req = [((0, ()),) + tuple((0, (i,)) for i in range(xd)) + tuple((0, (i, i)) for i in range(xd))]
    ep = epilogue.Epilogue(prob, eng.all_params, x, [(0, x.shape[0])], [[x]], req, net)
    assert ep.affine is not None
    A, B = ep.affine
    wl = plan.wl
    ops.forward(net.as_list(), eng.theta, plan.dom, plan.x, wl.tensor_list(), wl.size_list(), plan.pairbuf)
    q = torch.empty((net.J, x.shape[0]), device="cuda")
    ops.pou_reduce(net.as_list(), wl.tensor_list(), wl.size_list(), plan.pairbuf, plan.wsum, q)
    C = torch.empty_like(q)
    ops.pou_reduce(net.as_list(), wl.tensor_list(), wl.size_list(), plan.pairbuf, plan.wsum, C, A, B)
    cols = []   # the request code of every jet column, in column order
    for c in range(net.J):
        code = 0 if c == 0 else None
        for i in range(xd):
            if c == net.col1(i):
                code = 1 + i
            if (net.d2mask >> i) & 1 and c == net.col2(i):
                code = 1 + xd + i
        cols.append(code)
    outs = [torch.empty((x.shape[0], 1), device="cuda") for _ in cols]
    ops.affine_jets(net.as_list(), 0, x.shape[0], A, B, q, cols, outs)
    for c, o in enumerate(outs):
        assert torch.equal(C[c], o.reshape(-1)), c
    # backward
    gC = torch.randn(C.shape, generator=torch.Generator().manual_seed(1)).cuda()
    gseg_a = torch.empty_like(plan.gseg)
    ops.pou_backward(net.as_list(), wl.tensor_list(), wl.size_list(), plan.wsum, gC, gseg_a, B)
    gq = torch.empty_like(q)
    ops.affine_jets_backward(net.as_list(), 0, x.shape[0], B, cols, [gC[c].reshape(-1, 1).contiguous()
                                                                     for c in range(net.J)], gq)
    gseg_b = torch.empty_like(plan.gseg)
    ops.pou_backward(net.as_list(), wl.tensor_list(), wl.size_list(), plan.wsum, gq, gseg_b)
    e = ((gseg_a - gseg_b).abs().max() / gseg_b.abs().max()).item()
    print(f"  forward bit-exact, backward rel {e:.1e}", flush=True)
    assert e < 1e-6
    # g_scale multiplies the jets' gradient as it is read: the same as
    # scaling it beforehand, bit for bit (1 leaves it unchanged)
    for sv in (1.0, -2.75):
        s = torch.full((), sv, device="cuda")
        gseg_s = torch.empty_like(plan.gseg)
        ops.pou_backward(net.as_list(), wl.tensor_list(), wl.size_list(), plan.wsum, gC, gseg_s, B, s)
        gseg_r = torch.empty_like(plan.gseg)
        ops.pou_backward(net.as_list(), wl.tensor_list(), wl.size_list(), plan.wsum, gC * sv, gseg_r, B)
        assert torch.equal(gseg_s, gseg_r), sv
    # the POU transpose and a gathered vector's segment sum in one launch:
    # each output equals its own call's bit for bit, with and without g_scale
    g = torch.Generator().manual_seed(2)
    n_vec = 37
    idx = torch.randint(0, n_vec, (x.shape[0],), generator=g)
    perm = torch.argsort(idx, stable=True).to(torch.int32).cuda()
    seg_start = torch.zeros(n_vec + 1, dtype=torch.int32)
    seg_start[1:] = torch.cumsum(torch.bincount(idx, minlength=n_vec), 0)
    seg_start = seg_start.cuda()
    gv = torch.randn(x.shape[0], generator=g).cuda()
    for s in (None, torch.full((), -2.75, device="cuda")):
        gseg_1, out_1 = torch.empty_like(plan.gseg), torch.empty(n_vec, device="cuda")
        ops.pou_backward(net.as_list(), wl.tensor_list(), wl.size_list(), plan.wsum, gC, gseg_1, B, s)
        ops.segment_sum(seg_start, perm, gv, out_1, s)
        gseg_2, out_2 = torch.full_like(plan.gseg, float("nan")), torch.full((n_vec,), float("nan"), device="cuda")
        ops.pou_backward_segment_sum(net.as_list(), wl.tensor_list(), wl.size_list(), plan.wsum, gC, gseg_2, B, s,
                                     seg_start, perm, gv, out_2)
        assert torch.equal(gseg_1, gseg_2) and torch.equal(out_1, out_2), s
