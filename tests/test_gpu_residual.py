"""GPU: the fused pointwise-residual loss (fbpx_residual_loss and its
transpose, ops.ResidualLoss) against the problems' own loss_fn.

* every built-in problem with a residual description: the engine's loss
  epilogue with the fused kernels against the same epilogue running loss_fn,
  on the same random jets (float32 both): loss, gradient with respect to the
  jets of u~ and to trainable problem parameters within 1e-5 relative;
* sizes around the reduction's block and grid edges up to 2^20 + 37 points:
  the fused loss and its column gradients against loss_fn evaluated in
  float64 on the CPU, 1e-5 relative;
* the kernel's own term forms (constant term, a column squared, a column in
  two terms, a column without gradient) against the float64 formula;
* the per-point gather of a trainable vector and its transpose
  (fbpx_segment_sum) against float64 index_add, with unsorted indices and
  empty segments;
* two gathered columns with index vectors of their own, read as a term's
  second factor, as both factors of one term and beside plain columns: the
  forward with its column gradients and the backward with the coefficient
  gradients against float64;
* fbpx_residual_loss_grad (the loss launch that also writes the column
  gradients at unit upstream gradient) against fbpx_residual_loss followed
  by fbpx_residual_loss_backward with g_loss = 1, bit for bit, and
  fbpx_segment_sum's scale against scaling afterwards;
* fbpx_residual_losses (several residuals in one launch) against one
  launch per residual, bit for bit.
"""

import pytest
import torch

import setups
from fbpinns_amd import epilogue
from fbpinns_amd import problems as P
from fbpinns_amd.engine import pick_schema
from fbpinns_amd.ops import GatherRows, NetSpec, ResidualLoss

pytestmark = pytest.mark.gpu

PHYS1 = ((0, ()), (0, (0,)), (0, (0, 0)))
CASES = {
    "ho_soft": (P.HarmonicOscillator1D, dict(d=2, w0=20), 1, [PHYS1, ((0, ()), (0, (0,)))], [0, 2]),
    "ho_hard": (P.HarmonicOscillator1DHardBC, dict(d=2, w0=20), 1, [PHYS1], [0]),
    "ho_inverse": (P.HarmonicOscillator1DInverse, dict(d=2, w0=20), 1, [PHYS1, ((0, ()),)], [0, 1]),
    "burgers": (P.BurgersEquation2D, dict(), 2, [((0, ()), (0, (0,)), (0, (1,)), (0, (0, 0)))], [0]),
    "wave": (P.WaveEquationConstantVelocity3D, dict(c0=1.3, source=P.wave_sources()), 3,
             [((0, (0, 0)), (0, (1, 1)), (0, (2, 2)))], [0]),
    "wave_gauss": (setups.VariableVelocityWave3D, dict(c0=1.0, source=P.wave_sources()), 3,
                   [((0, (0, 0)), (0, (1, 1)), (0, (2, 2)))], [0]),
    "poisson": (P.PoissonMultiscale2D, dict(), 2, [((0, (0, 0)), (0, (1, 1)))], [0]),
    "batched_inverse": (P.BatchedHarmonicOscillatorInverse,
                        dict(w0s=[20.0, 35.0, 50.0, 80.0, 90.0], mus_true=[1.0, 2.0, 4.0, 6.0, 3.0], n_obs=4), 1,
                        [PHYS1, ((0, ()),)], [["i"], ["f", "i"]]),
}
N_INST = 5


def _value(kind, n, g):
    if kind == "i":
        return torch.randint(0, N_INST, (n,), generator=g)
    return torch.randn((n, 1), generator=g)


def _params(prob, kw, dev, trainable_value=0.7):
    s, t = prob.init_params(**kw)
    s = {k: (v.to(dev) if isinstance(v, torch.Tensor) else v) for k, v in s.items()}
    t = {k: (trainable_value + 0.1 * torch.arange(v.numel(), dtype=v.dtype).reshape(v.shape)).to(dev)
         .requires_grad_(True) for k, v in t.items()}
    return {"static": {"problem": s}, "trainable": {"problem": t}}


def _setup(name, sizes, dev, seed=0):
    prob, kw, xd, reqs, nvals = CASES[name]
    g = torch.Generator().manual_seed(seed)
    cranges, off = [], 0
    for s in sizes:
        cranges.append((off, s))
        off += s
    x = (torch.rand((off, xd), generator=g) * 0.8 + 0.1).to(dev)
    cvalues = []
    for (o, s), nv in zip(cranges, nvals):
        kinds = ["f"] * nv if isinstance(nv, int) else nv
        cvalues.append([x[o:o + s]] + [_value(k, s, g).to(dev) for k in kinds])
    d1, d2 = pick_schema(xd, *epilogue.schema_masks(reqs, xd))
    net = NetSpec(xd, 32, 2, d1, d2)
    return prob, kw, x, cranges, cvalues, reqs, net


@pytest.mark.parametrize("name", list(CASES))
def test_fused_epilogue_matches_loss_fn(name):
    dev = torch.device("cuda")
    sizes = [300, 1 if name == "ho_soft" else 13][:len(CASES[name][3])]
    prob, kw, x, cranges, cvalues, reqs, net = _setup(name, sizes, dev)
    ap = _params(prob, kw, dev)
    ef = epilogue.Epilogue(prob, ap, x, cranges, cvalues, reqs, net)
    eg = epilogue.Epilogue(prob, ap, x, cranges, cvalues, reqs, net, fused_loss=False)
    assert ef.fused is not None and eg.fused is None
    print(f"[residual] {name} residuals={len(ef.fused)} affine={ef.affine is not None}", flush=True)
    q = (torch.randn((net.J, x.shape[0]), generator=torch.Generator().manual_seed(1)) * 0.3).to(dev)
    q.requires_grad_(True)
    leaves = [q] + list(ap["trainable"]["problem"].values())
    lf = ef.loss(q, ap)
    lg = eg.loss(q, ap)
    e = abs(lf.item() - lg.item()) / abs(lg.item())
    gf = torch.autograd.grad(lf, leaves)
    gg = torch.autograd.grad(lg, leaves)
    print(f"  loss fused {lf.item():.6e} loss_fn {lg.item():.6e} rel {e:.1e}", flush=True)
    assert e < 1e-5
    for a, b in zip(gf, gg):
        eg_ = ((a - b).abs().max() / b.abs().max().clamp_min(1e-30)).item()
        print(f"  grad {tuple(b.shape)} rel {eg_:.1e}", flush=True)
        assert eg_ < 1e-5


@pytest.mark.parametrize("n", [1, 255, 256, 257, 65537, 1 << 20 | 37])
def test_fused_loss_sizes_against_float64(n):
    dev = torch.device("cuda")
    prob, kw, x, cranges, cvalues, reqs, net = _setup("wave_gauss", [n], dev, seed=n)
    ap = _params(prob, kw, dev)
    ef = epilogue.Epilogue(prob, ap, x, cranges, cvalues, reqs, net)
    assert ef.fused is not None
    print(f"[residual] n={n}", flush=True)
    g = torch.Generator().manual_seed(7)
    jets64 = [torch.randn((n, 1), generator=g, dtype=torch.float64) for _ in reqs[0]]
    jets = [j.float().to(dev).requires_grad_(True) for j in jets64]
    lf = ef._fused_loss([jets], ap)
    gf = torch.autograd.grad(lf, jets)
    j64 = [j.clone().requires_grad_(True) for j in jets64]
    ap64 = {"static": {"problem": {k: (v.cpu().double() if isinstance(v, torch.Tensor) else v)
                                   for k, v in ap["static"]["problem"].items()}}, "trainable": {"problem": {}}}
    l64 = prob.loss_fn(ap64, [[x.cpu().double()] + j64])
    g64 = torch.autograd.grad(l64, j64)
    e = abs(lf.item() - l64.item()) / abs(l64.item())
    print(f"  loss rel {e:.1e}", flush=True)
    assert e < 1e-5
    for a, b in zip(gf, g64):
        eg_ = ((a.cpu().double() - b).abs().max() / b.abs().max()).item()
        assert eg_ < 1e-5, eg_


def test_residual_term_forms():
    """r = 2 a^2 - 0.5 a b + 3 + b c, loss = 0.25 mean r^2, c without gradient."""
    dev = torch.device("cuda")
    n = 5000
    g = torch.Generator().manual_seed(3)
    cols64 = [torch.randn(n, generator=g, dtype=torch.float64) for _ in range(3)]
    coef64 = torch.tensor([2.0, -0.5, 3.0, 1.0], dtype=torch.float64)
    fa, fb = [0, 0, -1, 1], [0, 1, -1, 2]
    cols = [c.float().to(dev) for c in cols64]
    cols[0].requires_grad_(True)
    cols[1].requires_grad_(True)
    coef = coef64.float().to(dev).requires_grad_(True)
    done = torch.zeros(1, dtype=torch.int32, device=dev)
    lf = ResidualLoss.apply(coef, 0.25, fa, fb, done, *cols)
    ga, gb, gc = torch.autograd.grad(lf, [cols[0], cols[1], coef])
    c64 = [c.clone().requires_grad_(i < 2) for i, c in enumerate(cols64)]
    k64 = coef64.clone().requires_grad_(True)
    a, b, c = c64
    r = k64[0] * a * a + k64[1] * a * b + k64[2] + k64[3] * b * c
    l64 = 0.25 * (r ** 2).mean()
    ra, rb, rc = torch.autograd.grad(l64, [a, b, k64])
    assert abs(lf.item() - l64.item()) / l64.item() < 1e-5
    for u, v in ((ga, ra), (gb, rb), (gc, rc)):
        assert ((u.cpu().double() - v).abs().max() / v.abs().max()).item() < 1e-5
    assert int(done.item()) == 0   # the single-launch counter is left at 0
    # a second call reuses the counter and gives the same loss, bit for bit
    lf2 = ResidualLoss.apply(coef, 0.25, fa, fb, done, *cols)
    assert lf2.item() == lf.item() and int(done.item()) == 0


@pytest.mark.parametrize("n,n_vec", [(1, 1), (1000, 7), (70000, 300)])
def test_gather_transpose_matches_index_add(n, n_vec):
    dev = torch.device("cuda")
    g = torch.Generator().manual_seed(n)
    idx = torch.randint(0, max(n_vec - 2, 1), (n,), generator=g)      # the last indices stay empty
    v64 = torch.randn(n_vec, generator=g, dtype=torch.float64)
    gy64 = torch.randn(n, generator=g, dtype=torch.float64)
    perm = torch.argsort(idx, stable=True).to(torch.int32).to(dev)
    seg = torch.zeros(n_vec + 1, dtype=torch.int64)
    seg[1:] = torch.cumsum(torch.bincount(idx, minlength=n_vec), 0)
    v = v64.float().to(dev).requires_grad_(True)
    y = GatherRows.apply(v, idx.to(torch.int32).to(dev), perm, seg.to(torch.int32).to(dev))
    assert torch.equal(y.detach().cpu(), v64.float()[idx])
    (gv,) = torch.autograd.grad(y, v, gy64.float().to(dev))
    ref = torch.zeros(n_vec, dtype=torch.float64).index_add_(0, idx, gy64)
    e = ((gv.cpu().double() - ref).abs().max() / ref.abs().max().clamp_min(1e-30)).item()
    print(f"[gather] n={n} n_vec={n_vec} rel {e:.1e}", flush=True)
    assert e < 1e-5
    assert (gv[n_vec - 2:].cpu() == 0).all() if n_vec > 2 else True


@pytest.mark.parametrize("n", [200_000, 1000])
def test_single_launch_sum_equals_two_launch_sum(n):
    """fbpx_residual_loss with a workgroup counter (the last workgroup sums
    the partials) and without one (a second launch sums them): the same
    fixed-order sum, bit for bit, over 196 workgroups (the single-launch
    form is taken up to 256) and over one (whose total is the sum)."""
    import ctypes
    from fbpinns_amd import _lib
    dev = torch.device("cuda")
    g = torch.Generator().manual_seed(5)
    cols = [torch.randn(n, generator=g).to(dev) for _ in range(2)]
    coef = torch.tensor([1.0, 0.5], device=dev)
    fa = (ctypes.c_int32 * 2)(0, 1)
    fb = (ctypes.c_int32 * 2)(-1, 0)
    ptrs = (ctypes.c_void_p * 2)(*[c.data_ptr() for c in cols])
    out = []
    for use_done in (False, True):
        r = torch.empty(n, device=dev)
        scratch = torch.empty(_lib.FBPX_RESIDUAL_BLOCKS, device=dev)
        loss = torch.empty((), device=dev)
        done = torch.zeros(1, dtype=torch.int32, device=dev)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = _lib.lib().fbpx_residual_loss(stream, n, 2, ptrs, None, 2, fa, fb, ctypes.c_void_p(coef.data_ptr()),
                                           0.7, ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(scratch.data_ptr()),
                                           ctypes.c_void_p(loss.data_ptr()), 0,
                                           ctypes.c_void_p(done.data_ptr()) if use_done else None)
        assert rc == 0
        torch.cuda.synchronize()
        out.append(loss.item())
        assert int(done.item()) == 0
    ref = 0.7 * ((cols[0].double() + 0.5 * cols[1].double() * cols[0].double()) ** 2).mean().item()
    print(f"[residual] two launches {out[0]!r} one launch {out[1]!r} float64 {ref!r}", flush=True)
    assert out[0] == out[1]
    assert abs(out[0] - ref) / ref < 1e-5


@pytest.mark.parametrize("n,n_vec", [(1000, 7), (70_001, 3), (300_000, 256)])
def test_gathered_factors_in_every_position(n, n_vec):
    """Two columns gathered through two index vectors, read as a term's
    second factor, as both factors of one term and beside a plain column:
    r = c0 a v[i] + c1 v[i]^2 + c2 w[j] + c3 b w[j].  The forward with its
    column gradients (fbpx_residual_loss_grad) and the backward with the
    coefficient gradients (fbpx_residual_loss_backward) against float64
    torch: r and the loss 1e-5 relative, the gradients 1e-4."""
    from fbpinns_amd import _lib, ops
    dev = torch.device("cuda")
    g = torch.Generator().manual_seed(int(n) % 89 + 5)
    n_w = max(1, n_vec // 3)
    a, b = torch.randn(n, generator=g).to(dev), torch.randn(n, generator=g).to(dev)
    v, w = torch.randn(n_vec, generator=g).to(dev), torch.randn(n_w, generator=g).to(dev)
    iv = torch.randint(0, n_vec, (n,), generator=g).to(dev, torch.int32)
    iw = torch.randint(0, n_w, (n,), generator=g).to(dev, torch.int32)
    coef = torch.tensor([0.7, -1.1, 0.4, 1.6], device=dev)
    fa, fb = [0, 1, 2, 3], [1, 1, -1, 2]
    cols, weight = [a, v, w, b], 0.8
    print(f"[gather positions] n={n} n_vec={n_vec} n_w={n_w}", flush=True)
    r = torch.full((n,), float("nan"), device=dev)
    loss = torch.empty((), device=dev)
    done = torch.zeros(1, dtype=torch.int32, device=dev)
    gf = [torch.full((n,), float("nan"), device=dev) for _ in range(4)]
    scratch = torch.empty(_lib.FBPX_RESIDUAL_BLOCKS * 4, device=dev)
    ops.residual_loss_grad(cols, fa, fb, coef, weight, r, scratch, loss, done, [1, 2], [iv, iw], False,
                           [0, 1, 2, 3], gf)
    g_loss = torch.tensor([1.3], device=dev)
    gb = [torch.full((n,), float("nan"), device=dev) for _ in range(4)]
    gcoef = torch.full((4,), float("nan"), device=dev)
    ops.residual_loss_backward(cols, fa, fb, coef, r, weight, g_loss, [0, 1, 2, 3], gb, scratch, gcoef, True,
                               done, [1, 2], [iv, iw])
    torch.cuda.synchronize()
    a64, b64 = a.double().cpu(), b.double().cpu()
    vi, wj = v.double().cpu()[iv.long().cpu()], w.double().cpu()[iw.long().cpu()]
    c = coef.double().cpu()
    terms = [c[0] * a64 * vi, c[1] * vi * vi, c[2] * wj, c[3] * b64 * wj]
    rr = sum(terms)
    lref = weight * (rr ** 2).mean().item()
    w2n = 2.0 * weight / n
    dcol = [c[0] * vi, c[0] * a64 + 2.0 * c[1] * vi, c[2] + c[3] * b64, c[3] * wj]
    assert torch.allclose(r.double().cpu(), rr, rtol=1e-5, atol=1e-5)
    assert abs(loss.item() - lref) <= 1e-5 * abs(lref)
    for f in range(4):
        for scale, got in ((1.0, gf[f]), (1.3, gb[f])):
            ref = scale * w2n * rr * dcol[f]
            e = ((got.double().cpu() - ref).abs().max() / ref.abs().max()).item()
            assert e <= 1e-4, (f, scale, e)
    cref = torch.stack([1.3 * w2n * (rr * t / c[k]).sum() for k, t in enumerate(terms)])
    e = ((gcoef.double().cpu() - cref).abs().max() / cref.abs().max()).item()
    print(f"  loss {loss.item():.9e} float64 {lref:.9e}; coefficient gradients rel {e:.1e}", flush=True)
    assert e <= 1e-4
    assert int(done.item()) == 0


@pytest.mark.parametrize("n,n_vec,use_done", [(1000, 7, True), (300_000, 256, False), (300_000, 256, True)])
def test_gathered_column_and_accumulate(n, n_vec, use_done):
    """A column read through a per-point index (F = v[idx[k]], the batched
    inverse oscillators' damping per instance) and two residuals summed into
    one loss scalar (accumulate): loss, r and the backward's per-point
    gradient of the gathered column against float64 torch, 1e-5 relative
    (gradient 1e-4); with and without the single-launch sum."""
    import ctypes
    from fbpinns_amd import _lib
    dev = torch.device("cuda")
    g = torch.Generator().manual_seed(int(n) % 97)
    a = torch.randn(n, generator=g).to(dev)
    b = torch.randn(n, generator=g).to(dev)
    v = torch.randn(n_vec, generator=g).to(dev)
    idx = torch.randint(0, n_vec, (n,), generator=g).to(dev, torch.int32)
    coef = torch.tensor([1.0, 2.0, -0.5], device=dev)
    # r = a + 2 v[idx] b - 0.5
    fa, fb = [0, 2, -1], [-1, 1, -1]
    cols = [a, b, v]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = _lib.lib()
    ptrs = (ctypes.c_void_p * 3)(*[c.data_ptr() for c in cols])
    iptrs = (ctypes.c_void_p * 3)(None, None, idx.data_ptr())
    cfa, cfb = (ctypes.c_int32 * 3)(*fa), (ctypes.c_int32 * 3)(*fb)
    r = torch.empty(n, device=dev)
    scratch = torch.empty(_lib.FBPX_RESIDUAL_BLOCKS * 3, device=dev)
    loss = torch.full((), 0.25, device=dev)       # a first residual's loss, accumulated into
    done = torch.zeros(1, dtype=torch.int32, device=dev)
    dptr = ctypes.c_void_p(done.data_ptr()) if use_done else None
    assert L.fbpx_residual_loss(stream, n, 3, ptrs, iptrs, 3, cfa, cfb, ctypes.c_void_p(coef.data_ptr()), 0.6,
                                ctypes.c_void_p(r.data_ptr()), ctypes.c_void_p(scratch.data_ptr()),
                                ctypes.c_void_p(loss.data_ptr()), 1, dptr) == 0
    g_loss = torch.tensor([1.3], device=dev)
    gv = torch.empty(n, device=dev)
    gptrs = (ctypes.c_void_p * 3)(None, None, gv.data_ptr())
    assert L.fbpx_residual_loss_backward(stream, n, 3, ptrs, iptrs, 3, cfa, cfb, ctypes.c_void_p(coef.data_ptr()),
                                         ctypes.c_void_p(r.data_ptr()), 0.6, ctypes.c_void_p(g_loss.data_ptr()),
                                         gptrs, ctypes.c_void_p(scratch.data_ptr()), None, dptr) == 0
    torch.cuda.synchronize()
    a64, b64, v64 = a.double().cpu(), b.double().cpu(), v.double().cpu()
    vi = v64[idx.long().cpu()]
    rr = a64 + 2.0 * vi * b64 - 0.5
    lref = 0.25 + 0.6 * (rr ** 2).mean().item()
    gref = 1.3 * 2 * 0.6 / n * rr * 2.0 * b64
    print(f"[gather] n={n} loss {loss.item():.9e} float64 {lref:.9e}", flush=True)
    assert abs(loss.item() - lref) <= 1e-5 * abs(lref)
    assert torch.allclose(r.double().cpu(), rr, rtol=1e-5, atol=1e-5)
    assert ((gv.double().cpu() - gref).abs().max() / gref.abs().max()).item() <= 1e-4
    if use_done:
        assert int(done.item()) == 0


@pytest.mark.parametrize("n,n_vec,use_done", [(1000, 7, True), (300_000, 256, False), (300_000, 256, True)])
def test_loss_grad_launch_equals_loss_then_transpose(n, n_vec, use_done):
    """fbpx_residual_loss_grad: r and the loss equal fbpx_residual_loss's and
    the column gradients (two plain columns and a gathered one) equal
    fbpx_residual_loss_backward's at g_loss = 1, bit for bit; the gathered
    column's per-point gradient summed by fbpx_segment_sum with scale s
    equals the plain sum times s, bit for bit, and the transpose at g_loss = s
    within 1e-6 relative."""
    import ctypes
    from fbpinns_amd import _lib, ops
    dev = torch.device("cuda")
    g = torch.Generator().manual_seed(int(n) % 89)
    a = torch.randn(n, generator=g).to(dev)
    b = torch.randn(n, generator=g).to(dev)
    v = torch.randn(n_vec, generator=g).to(dev)
    idx = torch.randint(0, n_vec, (n,), generator=g).to(dev, torch.int32)
    coef = torch.tensor([1.0, 2.0, -0.5, 0.75], device=dev)
    # r = a + 2 v[idx] b - 0.5 + 0.75 a^2
    fa, fb = [0, 2, -1, 0], [-1, 1, -1, 0]
    cols = [a, b, v]
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L = _lib.lib()
    ptrs = (ctypes.c_void_p * 3)(*[c.data_ptr() for c in cols])
    iptrs = (ctypes.c_void_p * 3)(None, None, idx.data_ptr())
    cfa, cfb = (ctypes.c_int32 * 4)(*fa), (ctypes.c_int32 * 4)(*fb)
    cp = ctypes.c_void_p(coef.data_ptr())
    done = torch.zeros(1, dtype=torch.int32, device=dev)
    dptr = ctypes.c_void_p(done.data_ptr()) if use_done else None
    outs = {}
    for form in ("grad", "loss"):
        r = torch.empty(n, device=dev)
        scratch = torch.empty(_lib.FBPX_RESIDUAL_BLOCKS * 4, device=dev)
        loss = torch.full((), 0.25, device=dev)
        gr = [torch.full((n,), float("nan"), device=dev) for _ in range(3)]
        gptrs = (ctypes.c_void_p * 3)(*[t.data_ptr() for t in gr])
        args = (stream, n, 3, ptrs, iptrs, 4, cfa, cfb, cp, 0.6, ctypes.c_void_p(r.data_ptr()),
                ctypes.c_void_p(scratch.data_ptr()), ctypes.c_void_p(loss.data_ptr()), 1, dptr)
        if form == "grad":
            assert L.fbpx_residual_loss_grad(*args, gptrs) == 0
        else:
            assert L.fbpx_residual_loss(*args) == 0
            one = torch.ones(1, device=dev)
            assert L.fbpx_residual_loss_backward(stream, n, 3, ptrs, iptrs, 4, cfa, cfb, cp,
                                                 ctypes.c_void_p(r.data_ptr()), 0.6,
                                                 ctypes.c_void_p(one.data_ptr()), gptrs,
                                                 ctypes.c_void_p(scratch.data_ptr()), None, dptr) == 0
        torch.cuda.synchronize()
        outs[form] = (r, loss, gr)
        if use_done:
            assert int(done.item()) == 0
    (r1, l1, g1), (r0, l0, g0) = outs["grad"], outs["loss"]
    print(f"[loss grad] n={n} loss {l1.item():.9e} / {l0.item():.9e}", flush=True)
    assert torch.equal(r1, r0) and torch.equal(l1, l0)
    for f in range(3):
        assert torch.equal(g1[f], g0[f]), f
    # the gathered column's transpose with a scale
    perm = torch.argsort(idx.long(), stable=True).to(torch.int32)
    seg = torch.zeros(n_vec + 1, dtype=torch.int64, device=dev)
    seg[1:] = torch.cumsum(torch.bincount(idx.long(), minlength=n_vec), 0)
    seg = seg.to(torch.int32)
    s = torch.full((), -1.3, device=dev)
    plain, scaled = torch.empty(n_vec, device=dev), torch.empty(n_vec, device=dev)
    ops.segment_sum(seg, perm, g1[2], plain)
    ops.segment_sum(seg, perm, g1[2], scaled, s)
    assert torch.equal(scaled, plain * -1.3)
    gs = torch.empty(n, device=dev)
    scratch = torch.empty(_lib.FBPX_RESIDUAL_BLOCKS * 4, device=dev)
    sp = torch.full((1,), -1.3, device=dev)
    gptrs = (ctypes.c_void_p * 3)(None, None, gs.data_ptr())
    assert L.fbpx_residual_loss_backward(stream, n, 3, ptrs, iptrs, 4, cfa, cfb, cp, ctypes.c_void_p(r0.data_ptr()),
                                         0.6, ctypes.c_void_p(sp.data_ptr()), gptrs,
                                         ctypes.c_void_p(scratch.data_ptr()), None, dptr) == 0
    ref = torch.empty(n_vec, device=dev)
    ops.segment_sum(seg, perm, gs, ref)
    torch.cuda.synchronize()
    assert ((scaled - ref).abs().max() / ref.abs().max()).item() < 1e-6


# (points, terms) per residual.  Workgroups of the one launch: 1 + 1 and
# 50 + 16 (the last workgroup sums the partials), 1,024 + 293 + 69 (a second
# launch sums them), 128 + 128 and 128 + 129 (either side of the 256 where the
# second launch takes over), one residual alone; term counts reach each
# compiled maximum (2, 4, 8).
RESIDUAL_SETS = {
    "two_small": [(1000, 2), (300, 1)],
    "inverse_sizes": [(51_200, 3), (16_384, 2)],
    "three_large": [(1_000_003, 4), (300_000, 1), (70_001, 7)],
    "eight_terms": [(5000, 8), (257, 5), (1, 3), (4096, 2)],
    "grid_256": [(131_072, 2), (131_072, 3)],
    "grid_257": [(131_072, 2), (131_073, 3)],
    "one_residual": [(70_001, 3)],
}
_FA = [0, 2, -1, 0, 1, 2, 1, 0]
_FB = [-1, 1, -1, 0, 1, 0, -1, 2]


@pytest.mark.parametrize("acc", [False, True])
@pytest.mark.parametrize("grad", [True, False])
@pytest.mark.parametrize("name", list(RESIDUAL_SETS))
def test_residual_set_launch_equals_separate_launches(name, grad, acc):
    """fbpx_residual_losses: several residuals (two plain columns and one
    gathered through a per-point index, up to eight terms) in one launch give
    r, the column gradients at unit upstream gradient and the loss of
    fbpx_residual_loss_grad (or fbpx_residual_loss) launched once per
    residual, each adding into the loss, bit for bit."""
    from fbpinns_amd import _lib, ops
    dev = torch.device("cuda")
    g = torch.Generator().manual_seed(len(name) * 7 + grad + 2 * acc)
    spec = RESIDUAL_SETS[name]
    print(f"[residual set] {name} grad={grad} acc={acc}: {spec}", flush=True)
    data = []
    for n, nt in spec:
        n_vec = max(1, n // 200)
        cols = [torch.randn(n, generator=g).to(dev), torch.randn(n, generator=g).to(dev),
                torch.randn(n_vec, generator=g).to(dev)]
        idx = torch.randint(0, n_vec, (n,), generator=g).to(dev, torch.int32)
        coef = torch.randn(nt, generator=g).to(dev)
        weight = float(torch.rand((), generator=g)) + 0.5
        data.append((n, nt, cols, idx, coef, weight))
    outs = {}
    for form in ("set", "separate"):
        loss = torch.full((), 0.25, device=dev)
        done = torch.zeros(1, dtype=torch.int32, device=dev)
        rs, gcs, residuals = [], [], []
        for n, nt, cols, idx, coef, weight in data:
            rs.append(torch.full((n,), float("nan"), device=dev))
            gcs.append([torch.full((n,), float("nan"), device=dev), None,
                        torch.full((n,), float("nan"), device=dev)] if grad else [None] * 3)
            residuals.append(dict(n=n, weight=weight, coef=coef, fa=_FA[:nt], fb=_FB[:nt], cols=cols,
                                  idx=[None, None, idx], gcols=gcs[-1], r=rs[-1]))
        if form == "set":
            scratch = torch.empty(len(spec) * _lib.FBPX_RESIDUAL_BLOCKS, device=dev)
            ops.residual_losses(residuals, scratch, loss, done, accumulate=acc)
        else:
            for i, R in enumerate(residuals):
                scratch = torch.empty(_lib.FBPX_RESIDUAL_BLOCKS, device=dev)
                args = (R["cols"], R["fa"], R["fb"], R["coef"], R["weight"], R["r"], scratch, loss, done, [2],
                        [R["idx"][2]], acc or i > 0)
                if grad:
                    ops.residual_loss_grad(*args, [0, 2], [R["gcols"][0], R["gcols"][2]])
                else:
                    ops.residual_loss(*args)
        torch.cuda.synchronize()
        assert int(done.item()) == 0
        outs[form] = (loss, rs, gcs)
    (l1, r1, g1), (l0, r0, g0) = outs["set"], outs["separate"]
    print(f"  loss {l1.item():.9e} / {l0.item():.9e}", flush=True)
    assert torch.equal(l1, l0)
    for i in range(len(spec)):
        assert torch.equal(r1[i], r0[i]), i
        for a, b in zip(g1[i], g0[i]):
            assert (a is None) == (b is None) and (a is None or torch.equal(a, b)), i
