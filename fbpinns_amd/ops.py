"""Torch-facing operators over the C ABI of libfbpx.so.

Every operator here is a torch custom op (`torch.ops.fbpx.*`) whose body
validates device/dtype/contiguity and calls exactly one C-ABI entry point on
torch's current stream.  Nothing falls back to another implementation: if the
library is missing or a call fails, FbpxError is raised.

FBPINNJets is the torch.autograd.Function of the path: forward = pair jets +
sorted POU reduction, backward = POU transpose + the hand-written reverse
kernel (activation recompute) + deterministic per-subdomain row reduction.
"""

import ctypes
from typing import List, Optional

import torch

from fbpinns_amd import _lib
from fbpinns_amd._lib import FBPX_CHUNK, FBPX_DOM_STRIDE, FbpxError, check, lib

__all__ = ["FbpxError", "NetSpec", "Worklist", "FBPINNJets", "AffineJets", "FBPX_CHUNK", "FBPX_DOM_STRIDE"]


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t, dtype=None, name="tensor"):
    if t is None:
        return None
    if not t.is_cuda:
        raise FbpxError(f"{name} must be a GPU tensor")
    if dtype is not None and t.dtype != dtype:
        raise FbpxError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise FbpxError(f"{name} must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


class NetSpec:
    """Network [xd, h x nh, 1] and jet schema (d1mask, d2mask)."""

    def __init__(self, xd, h, nh, d1mask=0, d2mask=0):
        self.xd, self.h, self.nh, self.d1mask, self.d2mask = int(xd), int(h), int(nh), int(d1mask), int(d2mask)
        self.J = 1 + bin(self.d1mask).count("1") + bin(self.d2mask).count("1")
        self.P = _lib.param_count(self.xd, self.h, self.nh)

    def as_list(self):
        return [self.xd, self.h, self.nh, self.d1mask, self.d2mask]

    def c(self):
        return _lib.fbpx_net(*self.as_list())

    def with_schema(self, d1mask, d2mask):
        return NetSpec(self.xd, self.h, self.nh, d1mask, d2mask)

    def supported(self):
        return _lib.supported(*self.as_list())

    def occupancy(self):
        """(forward, backward) resident 256-thread workgroups per CU of the
        pair kernels for this shape (fbpx_occupancy; needs the GPU)."""
        f, b = ctypes.c_int32(0), ctypes.c_int32(0)
        check(lib().fbpx_occupancy(ctypes.byref(self.c()), ctypes.byref(f), ctypes.byref(b)), "fbpx_occupancy")
        return int(f.value), int(b.value)

    def col1(self, i):
        return 1 + bin(self.d1mask & ((1 << i) - 1)).count("1")

    def col2(self, i):
        return 1 + bin(self.d1mask).count("1") + bin(self.d2mask & ((1 << i) - 1)).count("1")

    def __repr__(self):
        return f"NetSpec(xd={self.xd}, h={self.h}, nh={self.nh}, d1mask={self.d1mask}, d2mask={self.d2mask})"


def _net_from_list(net: List[int]):
    return _lib.fbpx_net(*[int(v) for v in net])


# ------------------------------------------------------------------ work list

_WL_TENSORS = ["pair_point", "pair_seg", "chunk_model", "chunk_start", "point_seg_start", "seg_pair_start",
               "pair_slot", "fbin_chunk_start", "bbin_chunk_start"]
_WL_SIZES = ["n_pairs", "n_chunks", "n_points", "n_segs", "npou", "n_fbins", "n_bbins", "fwd_wave_bins",
             "bwd_wave_bins"]


class Worklist:
    """Device-resident work list (see fbpx_worklist in include/fbpx.h).

    tensors: dict of int32 CUDA tensors named as in _WL_TENSORS.
    sizes:   dict of ints named as in _WL_SIZES (the two wave flags default
             to 0: one workgroup per bin).
    """

    def __init__(self, tensors, sizes):
        self.t = tensors
        self.n = {k: int(v) for k, v in sizes.items()}
        for k in _WL_TENSORS:
            t = self.t[k]
            if t.dtype != torch.int32 or not t.is_contiguous():
                raise FbpxError(f"work list tensor {k} must be contiguous int32")

    def tensor_list(self):
        return [self.t[k] for k in _WL_TENSORS]

    def size_list(self):
        return [self.n.get(k, 0) for k in _WL_SIZES]

    def workgroups(self):
        """(forward, backward) workgroups launched: one per bin, or one per
        four bins with one wave per bin."""
        f, b = self.n["n_fbins"], self.n["n_bbins"]
        return ((f + 3) // 4 if self.n.get("fwd_wave_bins") else f,
                (b + 3) // 4 if self.n.get("bwd_wave_bins") else b)

    def with_wave_bins(self, fwd, bwd):
        """The same work list launched with one wave per forward / backward
        bin (fwd, bwd: bool; one-hidden-layer networks)."""
        return Worklist(self.t, dict(self.n, fwd_wave_bins=int(bool(fwd)), bwd_wave_bins=int(bool(bwd))))

    def c(self):
        return _wl_struct(self.tensor_list(), self.size_list())

    def __getattr__(self, k):
        n = self.__dict__.get("n", {})
        if k in n:
            return n[k]
        raise AttributeError(k)


def _wl_struct(ts: List[torch.Tensor], ns: List[int]):
    s = _lib.fbpx_worklist()
    for k, v in zip(_WL_SIZES, ns):
        setattr(s, k, int(v))
    for k, t in zip(_WL_TENSORS, ts):
        setattr(s, k, _p(t, torch.int32, k))
    return s


# ------------------------------------------------------------------ custom ops

@torch.library.custom_op("fbpx::inside_count", mutates_args=("counts",))
def inside_count(x: torch.Tensor, xmin: torch.Tensor, xmax: torch.Tensor, mask: torch.Tensor,
                 counts: torch.Tensor) -> None:
    n, xd = x.shape
    m = xmin.shape[0]
    mp = _p(mask, torch.uint8, "mask") if mask.numel() else None
    check(lib().fbpx_inside_count(_stream(), _p(x, torch.float32, "x"), n, xd, _p(xmin, torch.float32, "xmin"),
                                  _p(xmax, torch.float32, "xmax"), m, mp, _p(counts, torch.int32, "counts")),
          "fbpx_inside_count")


@torch.library.custom_op("fbpx::inside_fill", mutates_args=("n_take", "m_take"))
def inside_fill(x: torch.Tensor, xmin: torch.Tensor, xmax: torch.Tensor, mask: torch.Tensor,
                offsets: torch.Tensor, n_take: torch.Tensor, m_take: torch.Tensor) -> None:
    n, xd = x.shape
    m = xmin.shape[0]
    mp = _p(mask, torch.uint8, "mask") if mask.numel() else None
    check(lib().fbpx_inside_fill(_stream(), _p(x, torch.float32, "x"), n, xd, _p(xmin, torch.float32, "xmin"),
                                 _p(xmax, torch.float32, "xmax"), m, mp, _p(offsets, torch.int64, "offsets"),
                                 _p(n_take, torch.int32, "n_take"), _p(m_take, torch.int32, "m_take")),
          "fbpx_inside_fill")


@torch.library.custom_op("fbpx::pou_prepare", mutates_args=("scratch", "wsum"))
def pou_prepare(net: List[int], dom: torch.Tensor, x: torch.Tensor, wl_t: List[torch.Tensor], wl_n: List[int],
                scratch: torch.Tensor, wsum: torch.Tensor) -> None:
    n = _net_from_list(net)
    w = _wl_struct(wl_t, wl_n)
    check(lib().fbpx_pou_prepare(_stream(), ctypes.byref(n), _p(dom, torch.float32, "dom"),
                                 _p(x, torch.float32, "x"), ctypes.byref(w), _p(scratch, torch.float32, "scratch"),
                                 _p(wsum, torch.float32, "wsum")), "fbpx_pou_prepare")


@torch.library.custom_op("fbpx::forward", mutates_args=("pairbuf",))
def forward(net: List[int], theta: torch.Tensor, dom: torch.Tensor, x: torch.Tensor, wl_t: List[torch.Tensor],
            wl_n: List[int], pairbuf: torch.Tensor) -> None:
    n = _net_from_list(net)
    w = _wl_struct(wl_t, wl_n)
    check(lib().fbpx_forward(_stream(), ctypes.byref(n), _p(theta, torch.float32, "theta"),
                             _p(dom, torch.float32, "dom"), _p(x, torch.float32, "x"), ctypes.byref(w),
                             _p(pairbuf, torch.float32, "pairbuf")), "fbpx_forward")


@torch.library.custom_op("fbpx::pou_reduce", mutates_args=("ujet",))
def pou_reduce(net: List[int], wl_t: List[torch.Tensor], wl_n: List[int], pairbuf: torch.Tensor,
               wsum: torch.Tensor, ujet: torch.Tensor, A: Optional[torch.Tensor] = None,
               B: Optional[torch.Tensor] = None) -> None:
    """A, B: [J][n_points] jets of an affine constraining function (applied
    to the reduced jets), or None."""
    n = _net_from_list(net)
    w = _wl_struct(wl_t, wl_n)
    for t, nm in ((A, "A"), (B, "B")):
        if t is not None and t.shape != ujet.shape:
            raise FbpxError(f"pou_reduce: {nm} {tuple(t.shape)} for jets {tuple(ujet.shape)}")
    check(lib().fbpx_pou_reduce(_stream(), ctypes.byref(n), ctypes.byref(w), _p(pairbuf, torch.float32, "pairbuf"),
                                _p(wsum, torch.float32, "wsum"), _p(A, torch.float32, "A"), _p(B, torch.float32, "B"),
                                _p(ujet, torch.float32, "ujet")), "fbpx_pou_reduce")


@torch.library.custom_op("fbpx::backward_adam", mutates_args=("theta", "mu", "nu", "count", "count_done"))
def backward_adam(net: List[int], theta: torch.Tensor, dom: torch.Tensor, x: torch.Tensor, wl_t: List[torch.Tensor],
                  wl_n: List[int], chunk_train: torch.Tensor, gseg: torch.Tensor, g_scale: Optional[torch.Tensor],
                  mu: torch.Tensor, nu: torch.Tensor, count: torch.Tensor, count_done: torch.Tensor, lr: float,
                  b1: float, b2: float, eps: float) -> None:
    """fbpx_backward_adam: backward, then Adam on each trained subdomain row
    from its one chunk row, in the backward's own launch (one-hidden-layer
    networks, one chunk per trained row)."""
    n = _net_from_list(net)
    w = _wl_struct(wl_t, wl_n)
    if g_scale is not None and g_scale.numel() != 1:
        raise FbpxError(f"backward_adam: g_scale of {g_scale.numel()} values")
    if mu.shape != theta.shape or nu.shape != theta.shape:
        raise FbpxError(f"backward_adam: mu {tuple(mu.shape)} / nu {tuple(nu.shape)} for theta {tuple(theta.shape)}")
    check(lib().fbpx_backward_adam(_stream(), ctypes.byref(n), _p(theta, torch.float32, "theta"),
                                   _p(dom, torch.float32, "dom"), _p(x, torch.float32, "x"), ctypes.byref(w),
                                   _p(chunk_train, torch.uint8, "chunk_train"), _p(gseg, torch.float32, "gseg"),
                                   _p(g_scale, torch.float32, "g_scale"), _p(mu, torch.float32, "mu"),
                                   _p(nu, torch.float32, "nu"), _p(count, torch.int32, "count"),
                                   _p(count_done, torch.int32, "count_done"), float(lr), float(b1), float(b2),
                                   float(eps)),
          "fbpx_backward_adam")


@torch.library.custom_op("fbpx::pou_backward", mutates_args=("gseg",))
def pou_backward(net: List[int], wl_t: List[torch.Tensor], wl_n: List[int], wsum: torch.Tensor,
                 g_ujet: torch.Tensor, gseg: torch.Tensor, B: Optional[torch.Tensor] = None,
                 g_scale: Optional[torch.Tensor] = None) -> None:
    """g_scale: None, or a one-element tensor by which g_ujet is multiplied
    (d L / d loss for jets computed at unit upstream gradient)."""
    n = _net_from_list(net)
    w = _wl_struct(wl_t, wl_n)
    if B is not None and B.shape != g_ujet.shape:
        raise FbpxError(f"pou_backward: B {tuple(B.shape)} for jets {tuple(g_ujet.shape)}")
    if g_scale is not None and g_scale.numel() != 1:
        raise FbpxError(f"pou_backward: g_scale of {g_scale.numel()} values")
    check(lib().fbpx_pou_backward(_stream(), ctypes.byref(n), ctypes.byref(w), _p(wsum, torch.float32, "wsum"),
                                  _p(B, torch.float32, "B"), _p(g_ujet, torch.float32, "g_ujet"),
                                  _p(g_scale, torch.float32, "g_scale"), _p(gseg, torch.float32, "gseg")),
          "fbpx_pou_backward")


@torch.library.custom_op("fbpx::backward", mutates_args=("partial",))
def backward(net: List[int], theta: torch.Tensor, dom: torch.Tensor, x: torch.Tensor, wl_t: List[torch.Tensor],
             wl_n: List[int], chunk_train: torch.Tensor, gseg: torch.Tensor, partial: torch.Tensor,
             g_scale: Optional[torch.Tensor] = None) -> None:
    """g_scale: None, or a one-element tensor by which gseg is multiplied
    (d L / d loss for a gseg written at unit upstream gradient)."""
    n = _net_from_list(net)
    w = _wl_struct(wl_t, wl_n)
    ct = _p(chunk_train, torch.uint8, "chunk_train") if chunk_train.numel() else None
    if g_scale is not None and g_scale.numel() != 1:
        raise FbpxError(f"backward: g_scale of {g_scale.numel()} values")
    check(lib().fbpx_backward(_stream(), ctypes.byref(n), _p(theta, torch.float32, "theta"),
                              _p(dom, torch.float32, "dom"), _p(x, torch.float32, "x"), ctypes.byref(w), ct,
                              _p(gseg, torch.float32, "gseg"), _p(g_scale, torch.float32, "g_scale"),
                              _p(partial, torch.float32, "partial")),
          "fbpx_backward")


def _residual_array(residuals: list, name: str):
    """ctypes fbpx_residual[len(residuals)] for residual dicts: offset
    (default 0), n, weight, coef (tensor), fa, fb (term factor indices), cols
    (per column: an int jet column, or a tensor of per-point data / a gathered
    vector), idx (per column: None or the int32 gather index), gcols (per
    column: None or a per-point gradient tensor), r (the residual tensor)."""
    if not 1 <= len(residuals) <= _lib.FBPX_POU_LOSS_RES:
        raise FbpxError(f"{name}: {len(residuals)} residuals (at most {_lib.FBPX_POU_LOSS_RES})")
    arr = (_lib.fbpx_residual * len(residuals))()
    for d, R in zip(arr, residuals):
        cols, fa, fb = R["cols"], R["fa"], R["fb"]
        if not 1 <= len(cols) <= _lib.FBPX_POU_LOSS_COLS or not 1 <= len(fa) <= _lib.FBPX_MAX_TERMS \
                or len(fb) != len(fa):
            raise FbpxError(f"{name}: {len(cols)} columns (at most {_lib.FBPX_POU_LOSS_COLS}), {len(fa)}/{len(fb)} "
                            f"term factors (at most {_lib.FBPX_MAX_TERMS})")
        n = int(R["n"])
        if R["r"].numel() != n or R["coef"].numel() != len(fa):
            raise FbpxError(f"{name}: r of {R['r'].numel()} for {n} points, coef {R['coef'].numel()} for "
                            f"{len(fa)} terms")
        d.offset, d.n, d.n_cols, d.n_terms = int(R.get("offset", 0)), n, len(cols), len(fa)
        d.weight = float(R["weight"])
        d.coef = _p(R["coef"], torch.float32, "coef")
        d.r = _p(R["r"], torch.float32, "r")
        for t, (a, b) in enumerate(zip(fa, fb)):
            d.fa[t], d.fb[t] = int(a), int(b)
        for f, c in enumerate(cols):
            ix, gc = R["idx"][f], R["gcols"][f]
            if isinstance(c, int):
                d.jet[f] = c
            else:
                d.jet[f] = -1
                d.col[f] = _p(c, torch.float32, "column")
                if ix is None and c.numel() != n:
                    raise FbpxError(f"{name}: column of {c.numel()} values for {n} points")
            if ix is not None:
                if ix.numel() != n:
                    raise FbpxError(f"{name}: gather index of {ix.numel()} entries for {n} points")
                d.col_idx[f] = _p(ix, torch.int32, "gather index")
            if gc is not None:
                if gc.numel() != n:
                    raise FbpxError(f"{name}: gradient column of {gc.numel()} values for {n} points")
                d.gcol[f] = _p(gc, torch.float32, "gradient column")
    return arr


def residual_losses(residuals: list, scratch: torch.Tensor, loss: torch.Tensor, done: torch.Tensor,
                    accumulate: bool = False) -> None:
    """fbpx_residual_losses: several residual losses in one launch, each as
    residual_loss_grad (some gradient column given) or residual_loss would
    compute it, the loss summed as those launches one after another would.
    residuals: dicts as for pou_loss, every column a tensor.  scratch: at
    least len(residuals) * FBPX_RESIDUAL_BLOCKS floats.  A plain function
    for the same reason as pou_loss."""
    arr = _residual_array(residuals, "residual_losses")
    if loss.numel() != 1:
        raise FbpxError(f"residual_losses: loss of {loss.numel()} values")
    check(lib().fbpx_residual_losses(_stream(), len(residuals), arr, _p(scratch, torch.float32, "scratch"),
                                     scratch.numel(), _p(loss, torch.float32, "loss"), int(bool(accumulate)),
                                     _p(done, torch.int32, "done")),
          "fbpx_residual_losses")


def pou_loss(net: List[int], wl_t: List[torch.Tensor], wl_n: List[int], pairbuf: torch.Tensor, wsum: torch.Tensor,
             A: Optional[torch.Tensor], B: Optional[torch.Tensor], residuals: list, ujet: torch.Tensor,
             gseg: torch.Tensor, scratch: torch.Tensor, loss: torch.Tensor, done: torch.Tensor,
             accumulate: bool = False) -> None:
    """fbpx_pou_loss: POU reduction, affine constraint, the residual losses
    and the POU transpose at unit upstream gradient in one launch.
    residuals: dicts with offset, n, weight, coef (tensor), fa, fb (term
    factor indices), cols (per column: an int jet column, or a tensor of
    per-point data / a gathered vector), idx (per column: None or the int32
    gather index), gcols (per column: None or a per-point gradient tensor),
    r (the residual tensor).  The tensors stay referenced by the caller.
    A plain function rather than a torch custom op: its per-residual
    descriptors hold lists of tensors, and it is called only inside the
    fused loss's autograd node."""
    if (A is None) != (B is None) or (A is not None and (A.shape != ujet.shape or B.shape != ujet.shape)):
        raise FbpxError("pou_loss: A and B are both given (shaped as ujet) or both None")
    arr = _residual_array(residuals, "pou_loss")
    n_ = _net_from_list(net)
    w = _wl_struct(wl_t, wl_n)
    check(lib().fbpx_pou_loss(_stream(), ctypes.byref(n_), ctypes.byref(w), _p(pairbuf, torch.float32, "pairbuf"),
                              _p(wsum, torch.float32, "wsum"), _p(A, torch.float32, "A"), _p(B, torch.float32, "B"),
                              len(residuals), arr, _p(ujet, torch.float32, "ujet"), _p(gseg, torch.float32, "gseg"),
                              _p(scratch, torch.float32, "scratch"), scratch.numel(), _p(loss, torch.float32, "loss"),
                              int(bool(accumulate)), _p(done, torch.int32, "done")),
          "fbpx_pou_loss")


@torch.library.custom_op("fbpx::reduce_rows", mutates_args=("grad",))
def reduce_rows(partial: torch.Tensor, rows: torch.Tensor, row_chunk_start: torch.Tensor,
                grad: torch.Tensor) -> None:
    P = grad.shape[1]
    check(lib().fbpx_reduce_rows(_stream(), P, _p(partial, torch.float32, "partial"), _p(rows, torch.int32, "rows"),
                                 _p(row_chunk_start, torch.int32, "row_chunk_start"), rows.shape[0],
                                 _p(grad, torch.float32, "grad")), "fbpx_reduce_rows")


@torch.library.custom_op("fbpx::adam_rows", mutates_args=("theta", "mu", "nu", "count", "count_done"))
def adam_rows(theta: torch.Tensor, mu: torch.Tensor, nu: torch.Tensor, grad: torch.Tensor, rows: torch.Tensor,
              count: torch.Tensor, count_done: torch.Tensor, advance: bool, lr: float, b1: float, b2: float,
              eps: float) -> None:
    """advance: this launch also adds 1 to count (count_done: int32 0)."""
    P = theta.shape[1]
    check(lib().fbpx_adam_rows(_stream(), _p(theta, torch.float32, "theta"), _p(mu, torch.float32, "mu"),
                               _p(nu, torch.float32, "nu"), P, _p(grad, torch.float32, "grad"),
                               _p(rows, torch.int32, "rows"), rows.shape[0], _p(count, torch.int32, "count"),
                               _p(count_done, torch.int32, "count_done") if advance else None, lr, b1, b2, eps),
          "fbpx_adam_rows")


@torch.library.custom_op("fbpx::adam_rows_partial", mutates_args=("theta", "mu", "nu", "count", "count_done"))
def adam_rows_partial(theta: torch.Tensor, mu: torch.Tensor, nu: torch.Tensor, partial: torch.Tensor,
                      rows: torch.Tensor, row_chunk_start: torch.Tensor, count: torch.Tensor,
                      count_done: torch.Tensor, advance: bool, lr: float, b1: float, b2: float, eps: float) -> None:
    """reduce_rows + adam_rows in one launch (the gradient never materialised)."""
    P = theta.shape[1]
    if partial.shape[1] != P or row_chunk_start.numel() != rows.numel() + 1:
        raise FbpxError(f"adam_rows_partial: partial {tuple(partial.shape)}, {rows.numel()} rows, "
                        f"{row_chunk_start.numel()} row offsets")
    check(lib().fbpx_adam_rows_partial(_stream(), _p(theta, torch.float32, "theta"), _p(mu, torch.float32, "mu"),
                                       _p(nu, torch.float32, "nu"), P, _p(partial, torch.float32, "partial"),
                                       _p(rows, torch.int32, "rows"),
                                       _p(row_chunk_start, torch.int32, "row_chunk_start"), rows.shape[0],
                                       _p(count, torch.int32, "count"),
                                       _p(count_done, torch.int32, "count_done") if advance else None, lr, b1, b2,
                                       eps), "fbpx_adam_rows_partial")


@torch.library.custom_op("fbpx::adam_flat", mutates_args=("p", "mu", "nu", "count", "count_done"))
def adam_flat(p: torch.Tensor, mu: torch.Tensor, nu: torch.Tensor, g: torch.Tensor, count: torch.Tensor,
              count_done: torch.Tensor, advance: bool, lr: float, b1: float, b2: float, eps: float) -> None:
    check(lib().fbpx_adam_flat(_stream(), _p(p, torch.float32, "p"), _p(mu, torch.float32, "mu"),
                               _p(nu, torch.float32, "nu"), _p(g, torch.float32, "g"), p.numel(),
                               _p(count, torch.int32, "count"),
                               _p(count_done, torch.int32, "count_done") if advance else None, lr, b1, b2, eps),
          "fbpx_adam_flat")


@torch.library.custom_op("fbpx::count_increment", mutates_args=("count",))
def count_increment(count: torch.Tensor) -> None:
    check(lib().fbpx_count_increment(_stream(), _p(count, torch.int32, "count")), "fbpx_count_increment")


@torch.library.custom_op("fbpx::mean_into", mutates_args=("scratch", "out"))
def mean_into(x: torch.Tensor, scratch: torch.Tensor, out: torch.Tensor) -> None:
    check(lib().fbpx_mean(_stream(), _p(x, torch.float32, "x"), x.numel(), _p(scratch, torch.float32, "scratch"),
                          _p(out, torch.float32, "out")), "fbpx_mean")


def _jet_ptr(t, off, name):
    """Pointer to column 0, point `off` of a contiguous float32 [J][ld] jet array."""
    return ctypes.c_void_p(_p(t, torch.float32, name).value + 4 * int(off))


def _req_arrays(req, tensors, n, name):
    if not 1 <= len(req) <= _lib.FBPX_MAX_REQ or len(tensors) != len(req):
        raise FbpxError(f"{name}: {len(req)} requests ({len(tensors)} tensors), at most {_lib.FBPX_MAX_REQ}")
    for t in tensors:
        _p(t, torch.float32, name)
        if t.numel() < n:
            raise FbpxError(f"{name}: per-request tensor holds {t.numel()} < {n} values")
    codes = (ctypes.c_int32 * len(req))(*[int(c) for c in req])
    ptrs = (ctypes.c_void_p * len(req))(*[t.data_ptr() for t in tensors])
    return codes, ptrs


@torch.library.custom_op("fbpx::affine_jets", mutates_args=("outs",))
def affine_jets(net: List[int], off: int, n: int, A: torch.Tensor, B: torch.Tensor, q: torch.Tensor,
                req: List[int], outs: List[torch.Tensor]) -> None:
    ld = q.shape[1]
    if A.shape != q.shape or B.shape != q.shape or off < 0 or off + n > ld:
        raise FbpxError(f"affine_jets: jet arrays {tuple(A.shape)}, {tuple(B.shape)}, {tuple(q.shape)}, "
                        f"range [{off}, {off + n})")
    codes, ptrs = _req_arrays(req, outs, n, "affine_jets")
    nt = _net_from_list(net)
    check(lib().fbpx_affine_jets(_stream(), ctypes.byref(nt), n, ld, _jet_ptr(A, off, "A"), _jet_ptr(B, off, "B"),
                                 _jet_ptr(q, off, "q"), len(req), codes, ptrs), "fbpx_affine_jets")


@torch.library.custom_op("fbpx::affine_jets_backward", mutates_args=("gq",))
def affine_jets_backward(net: List[int], off: int, n: int, B: torch.Tensor, req: List[int],
                         gouts: List[torch.Tensor], gq: torch.Tensor) -> None:
    ld = gq.shape[1]
    if B.shape != gq.shape or off < 0 or off + n > ld:
        raise FbpxError(f"affine_jets_backward: jet arrays {tuple(B.shape)}, {tuple(gq.shape)}, "
                        f"range [{off}, {off + n})")
    codes, ptrs = _req_arrays(req, gouts, n, "affine_jets_backward")
    nt = _net_from_list(net)
    check(lib().fbpx_affine_jets_backward(_stream(), ctypes.byref(nt), n, ld, _jet_ptr(B, off, "B"), len(req),
                                          codes, ptrs, _jet_ptr(gq, off, "gq")), "fbpx_affine_jets_backward")


class AffineJets(torch.autograd.Function):
    """Requested derivatives of constraining_fn(x, u~) = A + B u~ for every
    constraint (Leibniz rule on jets, fbpx_affine_jets); one output tensor
    (n_c, 1) per request, in constraint order.  Gradient w.r.t. the jets of
    u~ only (A and B do not depend on trainable parameters)."""

    @staticmethod
    def forward(ctx, ujet, A, B, net, cranges, reqs):
        outs = []
        for (off, n_c), req in zip(cranges, reqs):
            o = [torch.empty((n_c, 1), dtype=torch.float32, device=ujet.device) for _ in req]
            if n_c:
                affine_jets(net, off, n_c, A, B, ujet, list(req), o)
            outs += o
        ctx.save_for_backward(B)
        ctx.net, ctx.cranges, ctx.reqs, ctx.shape = net, cranges, reqs, ujet.shape
        return tuple(outs)

    @staticmethod
    def backward(ctx, *gouts):
        (B,) = ctx.saved_tensors
        gq = torch.empty(ctx.shape, dtype=torch.float32, device=B.device)
        i = 0
        for (off, n_c), req in zip(ctx.cranges, ctx.reqs):
            g = [t.contiguous() for t in gouts[i:i + len(req)]]
            i += len(req)
            if n_c:
                affine_jets_backward(ctx.net, off, n_c, B, list(req), g, gq)
        return gq, None, None, None, None, None


def _residual_tables(cols, fa, fb, n, name, gathered=(), idx=()):
    """Host tables of one residual: column pointers, per-column gather index
    pointers (None when no column is gathered) and the factor indices.
    gathered[j] is the column whose values are cols[gathered[j]][idx[j][k]]."""
    if not 1 <= len(cols) <= _lib.FBPX_MAX_REQ or not 1 <= len(fa) <= _lib.FBPX_MAX_TERMS or len(fb) != len(fa):
        raise FbpxError(f"{name}: {len(cols)} columns (at most {_lib.FBPX_MAX_REQ}), {len(fa)}/{len(fb)} "
                        f"term factors (at most {_lib.FBPX_MAX_TERMS})")
    if len(gathered) != len(idx) or any(not 0 <= f < len(cols) for f in gathered):
        raise FbpxError(f"{name}: gathered columns {list(gathered)} with {len(idx)} index arrays")
    for f, t in enumerate(cols):
        _p(t, torch.float32, name)
        if f not in gathered and t.numel() != n:
            raise FbpxError(f"{name}: column of {t.numel()} values for {n} points")
    ptrs = (ctypes.c_void_p * len(cols))(*[t.data_ptr() for t in cols])
    iptrs = None
    if gathered:
        iptrs = (ctypes.c_void_p * len(cols))()
        for f, ix in zip(gathered, idx):
            _p(ix, torch.int32, f"{name} index")
            if ix.numel() != n:
                raise FbpxError(f"{name}: gather index of {ix.numel()} entries for {n} points")
            iptrs[f] = ix.data_ptr()
    a = (ctypes.c_int32 * len(fa))(*[int(v) for v in fa])
    b = (ctypes.c_int32 * len(fb))(*[int(v) for v in fb])
    return ptrs, iptrs, a, b


def _gradient_table(cols, gidx, gcols, n, name):
    """Host table of per-column gradient pointers: gcols[j] (n floats) for
    column gidx[j], NULL for the columns without a gradient."""
    if len(gidx) != len(gcols) or any(not 0 <= i < len(cols) for i in gidx):
        raise FbpxError(f"{name}: gradient columns {gidx} of {len(cols)}")
    gp = (ctypes.c_void_p * len(cols))()
    for i, g in zip(gidx, gcols):
        _p(g, torch.float32, "gcol")
        if g.numel() != n:
            raise FbpxError(f"{name}: gradient column of {g.numel()} values for {n} points")
        gp[i] = g.data_ptr()
    return gp


@torch.library.custom_op("fbpx::residual_loss", mutates_args=("r", "scratch", "loss", "done"))
def residual_loss(cols: List[torch.Tensor], fa: List[int], fb: List[int], coef: torch.Tensor, weight: float,
                  r: torch.Tensor, scratch: torch.Tensor, loss: torch.Tensor, done: torch.Tensor,
                  gathered: List[int], idx: List[torch.Tensor], accumulate: bool) -> None:
    """done: int32 0 (left 0), the single-launch workgroup counter.
    gathered / idx: columns read through a per-point index (their tensor is
    the vector gathered from); accumulate: add to loss instead of writing it."""
    n = r.numel()
    ptrs, iptrs, a, b = _residual_tables(cols, fa, fb, n, "residual_loss", gathered, idx)
    if coef.numel() != len(fa) or scratch.numel() < _lib.FBPX_RESIDUAL_BLOCKS or loss.numel() != 1:
        raise FbpxError(f"residual_loss: coef {coef.numel()} for {len(fa)} terms, scratch {scratch.numel()}, "
                        f"loss {loss.numel()}")
    check(lib().fbpx_residual_loss(_stream(), n, len(cols), ptrs, iptrs, len(fa), a, b,
                                   _p(coef, torch.float32, "coef"), float(weight), _p(r, torch.float32, "r"),
                                   _p(scratch, torch.float32, "scratch"), _p(loss, torch.float32, "loss"),
                                   int(bool(accumulate)), _p(done, torch.int32, "done")),
          "fbpx_residual_loss")


@torch.library.custom_op("fbpx::residual_loss_grad", mutates_args=("r", "scratch", "loss", "done", "gcols"))
def residual_loss_grad(cols: List[torch.Tensor], fa: List[int], fb: List[int], coef: torch.Tensor, weight: float,
                       r: torch.Tensor, scratch: torch.Tensor, loss: torch.Tensor, done: torch.Tensor,
                       gathered: List[int], idx: List[torch.Tensor], accumulate: bool, gidx: List[int],
                       gcols: List[torch.Tensor]) -> None:
    """residual_loss and, in the same launch, d loss / d cols[gidx[j]] at unit
    upstream gradient into gcols[j] (one value per point)."""
    n = r.numel()
    ptrs, iptrs, a, b = _residual_tables(cols, fa, fb, n, "residual_loss_grad", gathered, idx)
    if coef.numel() != len(fa) or scratch.numel() < _lib.FBPX_RESIDUAL_BLOCKS or loss.numel() != 1:
        raise FbpxError(f"residual_loss_grad: coef {coef.numel()} for {len(fa)} terms, scratch {scratch.numel()}, "
                        f"loss {loss.numel()}")
    gptrs = _gradient_table(cols, gidx, gcols, n, "residual_loss_grad")
    check(lib().fbpx_residual_loss_grad(_stream(), n, len(cols), ptrs, iptrs, len(fa), a, b,
                                        _p(coef, torch.float32, "coef"), float(weight), _p(r, torch.float32, "r"),
                                        _p(scratch, torch.float32, "scratch"), _p(loss, torch.float32, "loss"),
                                        int(bool(accumulate)), _p(done, torch.int32, "done"), gptrs),
          "fbpx_residual_loss_grad")


@torch.library.custom_op("fbpx::residual_loss_backward", mutates_args=("gcols", "scratch", "gcoef", "done"))
def residual_loss_backward(cols: List[torch.Tensor], fa: List[int], fb: List[int], coef: torch.Tensor,
                           r: torch.Tensor, weight: float, g_loss: torch.Tensor, gidx: List[int],
                           gcols: List[torch.Tensor], scratch: torch.Tensor, gcoef: torch.Tensor,
                           want_coef: bool, done: torch.Tensor, gathered: List[int], idx: List[torch.Tensor]) -> None:
    n = r.numel()
    ptrs, iptrs, a, b = _residual_tables(cols, fa, fb, n, "residual_loss_backward", gathered, idx)
    if coef.numel() != len(fa) or gcoef.numel() != len(fa) or scratch.numel() < _lib.FBPX_RESIDUAL_BLOCKS * len(fa):
        raise FbpxError(f"residual_loss_backward: coef {coef.numel()} / gcoef {gcoef.numel()} for {len(fa)} "
                        f"terms, scratch {scratch.numel()}")
    gp = _gradient_table(cols, gidx, gcols, n, "residual_loss_backward")
    check(lib().fbpx_residual_loss_backward(
        _stream(), n, len(cols), ptrs, iptrs, len(fa), a, b, _p(coef, torch.float32, "coef"),
        _p(r, torch.float32, "r"), float(weight), _p(g_loss, torch.float32, "g_loss"), gp,
        _p(scratch, torch.float32, "scratch"), _p(gcoef, torch.float32, "gcoef") if want_coef else None,
        _p(done, torch.int32, "done")),
        "fbpx_residual_loss_backward")


class ResidualLoss(torch.autograd.Function):
    """weight / n * sum_k (sum_t coef[t] F(fa[t], k) F(fb[t], k))^2 over n
    points (fbpx_residual_loss): cols are (n, 1) or (n,) float32 columns,
    F(f, k) = cols[f][k] (1 for f < 0), coef an (n_terms,) float32 tensor,
    done an int32 0 that the launches leave 0 (one per residual: residuals
    that may run concurrently need their own).  Gradients for the columns
    and coefficients that require them."""

    @staticmethod
    def forward(ctx, coef, weight, fa, fb, done, *cols):
        cs = [c.contiguous() for c in cols]
        dev = cs[0].device
        n = cs[0].numel()
        r = torch.empty(n, dtype=torch.float32, device=dev)
        scratch = torch.empty(_lib.FBPX_RESIDUAL_BLOCKS, dtype=torch.float32, device=dev)
        loss = torch.empty((), dtype=torch.float32, device=dev)
        residual_loss(cs, list(fa), list(fb), coef, float(weight), r, scratch, loss, done, [], [], False)
        ctx.save_for_backward(coef, r, done, *cs)
        ctx.weight, ctx.fa, ctx.fb = float(weight), list(fa), list(fb)
        return loss

    @staticmethod
    def backward(ctx, g):
        coef, r, done, *cs = ctx.saved_tensors
        need = ctx.needs_input_grad
        gidx = [i for i in range(len(cs)) if need[5 + i]]
        gcols = [torch.empty_like(cs[i]) for i in gidx]
        nt = len(ctx.fa)
        gcoef = torch.empty(nt, dtype=torch.float32, device=r.device)
        scratch = torch.empty(_lib.FBPX_RESIDUAL_BLOCKS * nt, dtype=torch.float32, device=r.device)
        residual_loss_backward(cs, ctx.fa, ctx.fb, coef, r, ctx.weight, g.contiguous(), gidx, gcols, scratch,
                               gcoef, bool(need[0]), done, [], [])
        out = [None] * len(cs)
        for i, gc in zip(gidx, gcols):
            out[i] = gc
        return (gcoef if need[0] else None, None, None, None, None, *out)


@torch.library.custom_op("fbpx::segment_sum", mutates_args=("out",))
def segment_sum(seg_start: torch.Tensor, perm: torch.Tensor, x: torch.Tensor, out: torch.Tensor,
                scale: Optional[torch.Tensor] = None) -> None:
    n_seg = out.numel()
    if seg_start.numel() != n_seg + 1 or perm.numel() != x.numel() or (scale is not None and scale.numel() != 1):
        raise FbpxError(f"segment_sum: {seg_start.numel()} offsets for {n_seg} segments, "
                        f"{perm.numel()} permutation entries for {x.numel()} values"
                        + ("" if scale is None else f", scale of {scale.numel()} values"))
    check(lib().fbpx_segment_sum(_stream(), n_seg, _p(seg_start, torch.int32, "seg_start"),
                                 _p(perm, torch.int32, "perm"), _p(x, torch.float32, "x"),
                                 _p(scale, torch.float32, "scale"), _p(out, torch.float32, "out")),
          "fbpx_segment_sum")


@torch.library.custom_op("fbpx::pou_backward_segment_sum", mutates_args=("gseg", "out"))
def pou_backward_segment_sum(net: List[int], wl_t: List[torch.Tensor], wl_n: List[int], wsum: torch.Tensor,
                             g_ujet: torch.Tensor, gseg: torch.Tensor, B: Optional[torch.Tensor],
                             g_scale: Optional[torch.Tensor], seg_start: torch.Tensor, perm: torch.Tensor,
                             x: torch.Tensor, out: torch.Tensor) -> None:
    """pou_backward(..., B, g_scale) and segment_sum(seg_start, perm, x, out,
    g_scale) in one launch, bit for bit."""
    n = _net_from_list(net)
    w = _wl_struct(wl_t, wl_n)
    n_seg = out.numel()
    if B is not None and B.shape != g_ujet.shape:
        raise FbpxError(f"pou_backward_segment_sum: B {tuple(B.shape)} for jets {tuple(g_ujet.shape)}")
    if seg_start.numel() != n_seg + 1 or perm.numel() != x.numel() or (g_scale is not None and g_scale.numel() != 1):
        raise FbpxError(f"pou_backward_segment_sum: {seg_start.numel()} offsets for {n_seg} segments, "
                        f"{perm.numel()} permutation entries for {x.numel()} values"
                        + ("" if g_scale is None else f", g_scale of {g_scale.numel()} values"))
    check(lib().fbpx_pou_backward_segment_sum(
        _stream(), ctypes.byref(n), ctypes.byref(w), _p(wsum, torch.float32, "wsum"), _p(B, torch.float32, "B"),
        _p(g_ujet, torch.float32, "g_ujet"), _p(g_scale, torch.float32, "g_scale"), _p(gseg, torch.float32, "gseg"),
        n_seg, _p(seg_start, torch.int32, "seg_start"), _p(perm, torch.int32, "perm"), _p(x, torch.float32, "x"),
        _p(out, torch.float32, "out")),
        "fbpx_pou_backward_segment_sum")


class GatherRows(torch.autograd.Function):
    """y[k] = v[idx[k]] for a trainable vector v; the backward sums the
    incoming gradient per index with fbpx_segment_sum over the points sorted
    by index (perm, seg_start: precomputed, the index is static)."""

    @staticmethod
    def forward(ctx, v, idx, perm, seg_start):
        ctx.save_for_backward(perm, seg_start)
        ctx.n = v.numel()
        return v.reshape(-1).index_select(0, idx)

    @staticmethod
    def backward(ctx, g):
        perm, seg_start = ctx.saved_tensors
        out = torch.empty(ctx.n, dtype=torch.float32, device=g.device)
        segment_sum(seg_start, perm, g.contiguous(), out)
        return out, None, None, None


class _Mean(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        xc = x.contiguous()
        out = torch.empty((), dtype=torch.float32, device=x.device)
        scratch = torch.empty(256, dtype=torch.float32, device=x.device)
        mean_into(xc, scratch, out)
        ctx.shape, ctx.n = x.shape, x.numel()
        return out

    @staticmethod
    def backward(ctx, g):
        return (g / ctx.n).expand(ctx.shape)


def mean(x):
    """Graph-safe deterministic mean (jnp.mean).  GPU float32 tensors run
    fbpx_mean; other tensors (CPU set-up work, float64 probes) use torch."""
    if x.is_cuda and x.dtype == torch.float32 and x.numel() > 0:
        return _Mean.apply(x)
    return torch.mean(x)


# ------------------------------------------------------------------ take builder

_EMPTY_U8 = {}


def _empty_u8(device):
    if device not in _EMPTY_U8:
        _EMPTY_U8[device] = torch.empty(0, dtype=torch.uint8, device=device)
    return _EMPTY_U8[device]


def inside_counts(x, xmins, xmaxs, mask=None):
    """Per-point number of boxes containing the point (GPU)."""
    counts = torch.empty(x.shape[0], dtype=torch.int32, device=x.device)
    inside_count(x, xmins, xmaxs, mask if mask is not None else _empty_u8(x.device), counts)
    return counts


def inside_takes(x, xmins, xmaxs, mask=None):
    """Row-major nonzero of the inside matrix on the GPU: (n_take, m_take)
    int32 CUDA tensors (decompositions_base.py:70-76)."""
    counts = inside_counts(x, xmins, xmaxs, mask)
    offsets = torch.zeros(x.shape[0] + 1, dtype=torch.int64, device=x.device)
    torch.cumsum(counts, 0, out=offsets[1:])
    s = int(offsets[-1].item())
    n_take = torch.empty(s, dtype=torch.int32, device=x.device)
    m_take = torch.empty(s, dtype=torch.int32, device=x.device)
    if s:
        inside_fill(x, xmins, xmaxs, mask if mask is not None else _empty_u8(x.device), offsets, n_take, m_take)
    return n_take, m_take, counts


def device_worklist(x, xmins, xmaxs, active, pou, chunk=0, fbins=0, bbins=0):
    """Takes and work list of points x (n, xd) for boxes xmins / xmaxs (m, xd)
    under the scheduler state `active` (m,) with levels `pou` (m,), built on
    the GPU by fbpx_takes_count + fbpx_build_worklist (get_inputs,
    trainers.py:332-419, and the fbpx_worklist layout).  chunk / fbins /
    bbins: 0 = none (see include/fbpx.h).  Returns (tensors, info): a dict of
    int32 CUDA tensors cut to their exact sizes and the info sizes as a dict."""
    dev = x.device
    x = x.to(torch.float32).contiguous()
    xmins = xmins.to(device=dev, dtype=torch.float32).contiguous()
    xmaxs = xmaxs.to(device=dev, dtype=torch.float32).contiguous()
    active = torch.as_tensor(active).to(device=dev, dtype=torch.int32).contiguous()
    pou = torch.as_tensor(pou).to(device=dev, dtype=torch.int32).reshape(-1).contiguous()
    n, xd = x.shape
    m = xmins.shape[0]
    if xmins.shape != (m, xd) or xmaxs.shape != (m, xd) or active.numel() != m or pou.numel() != m:
        raise FbpxError(f"device_worklist: shapes x {tuple(x.shape)} boxes {tuple(xmins.shape)} / "
                        f"{tuple(xmaxs.shape)} active {active.numel()} pou {pou.numel()}")
    L = lib()
    st = _stream()
    w0 = L.fbpx_takes_work_bytes(n, xd, m, 0, 0, 0)
    if w0 < 0:
        raise FbpxError(f"device_worklist: bad sizes n={n} xd={xd} m={m}")
    work = torch.empty(w0, dtype=torch.uint8, device=dev)
    offsets = torch.empty(n + 1, dtype=torch.int64, device=dev)
    s_c = ctypes.c_int64()
    check(L.fbpx_takes_count(st, _p(x, torch.float32, "x") if n else None, n, xd, xmins.data_ptr(),
                             xmaxs.data_ptr(), m, offsets.data_ptr(), work.data_ptr(), w0, ctypes.byref(s_c)),
          "fbpx_takes_count")
    s = int(s_c.value)
    max_chunks = int(L.fbpx_takes_max_chunks(m, s, chunk, fbins, bbins))
    wb = L.fbpx_takes_work_bytes(n, xd, m, s, fbins, bbins)
    del work
    work = torch.empty(wb, dtype=torch.uint8, device=dev)
    i32 = lambda k: torch.empty(max(int(k), 1), dtype=torch.int32, device=dev)
    t = {"m_take": i32(s), "n_take": i32(s), "p_take": i32(s), "np_take": i32(s), "active": i32(m),
         "all_ims": i32(m), "pair_point": i32(s), "pair_seg": i32(s), "pair_slot": i32(s),
         "point_seg_start": i32(n + 1), "seg_pair_start": i32(s + 1), "chunk_model": i32(max_chunks),
         "chunk_start": i32(max_chunks + 1),
         "chunk_train": torch.empty(max(max_chunks, 1), dtype=torch.uint8, device=dev),
         "row_chunk_start": i32(m + 1),
         "fbin_chunk_start": i32((fbins if fbins > 0 else max_chunks) + 1),
         "bbin_chunk_start": i32((bbins if bbins > 0 else max_chunks) + 1)}
    out = _lib.fbpx_takes_out(**{k: v.data_ptr() for k, v in t.items()})
    info = _lib.fbpx_takes_info()
    check(L.fbpx_build_worklist(st, x.data_ptr() if n else None, n, xd, xmins.data_ptr(), xmaxs.data_ptr(), m,
                                active.data_ptr(), pou.data_ptr(), offsets.data_ptr(), s, int(chunk), int(fbins),
                                int(bbins), max_chunks, work.data_ptr(), wb, ctypes.byref(out), ctypes.byref(info)),
          "fbpx_build_worklist")
    nf = {k: int(getattr(info, k)) for k, _ in _lib.fbpx_takes_info._fields_}
    cut = {"m_take": s, "n_take": s, "p_take": s, "np_take": nf["n_segs"], "active": m, "all_ims": nf["n_used"],
           "pair_point": s, "pair_seg": s, "pair_slot": s, "point_seg_start": n + 1,
           "seg_pair_start": nf["n_segs"] + 1, "chunk_model": nf["n_chunks"], "chunk_start": nf["n_chunks"] + 1,
           "chunk_train": nf["n_chunks"], "row_chunk_start": nf["n_active"] + 1,
           "fbin_chunk_start": nf["n_fbins"] + 1, "bbin_chunk_start": nf["n_bbins"] + 1}
    return {k: t[k][:cut[k]] for k in t}, nf


# ------------------------------------------------------------------ autograd


class Plan:
    """Everything the jets need for one set of takes (built by engine.build_plan)."""

    def __init__(self, net, dom, x, wl, wsum, chunk_train, rows, row_chunk_start, m):
        self.net, self.dom, self.x, self.wl, self.wsum = net, dom, x, wl, wsum
        self.chunk_train, self.rows, self.row_chunk_start, self.m = chunk_train, rows, row_chunk_start, m
        self.pairbuf = torch.empty((max(wl.n_pairs, 1), net.J), dtype=torch.float32, device=x.device)
        self.partial = torch.empty((max(wl.n_chunks, 1), net.P), dtype=torch.float32, device=x.device)
        self.gseg = torch.empty((max(wl.n_segs, 1), net.J), dtype=torch.float32, device=x.device)
        # set by the engine for its step plan: the backward returns this one
        # [m][P] buffer every time (untrained rows stay 0, so no zero fill per
        # step); callers that keep a gradient across steps copy it
        self.g_theta = None
        # set by the engine's step: the backward leaves the per-chunk rows in
        # `partial` for adam_rows_partial and returns no theta gradient
        self.defer_rows = False
        # set by the engine when every trained row is one chunk of a
        # one-hidden-layer network: (theta, mu, nu, count, count_done, lr,
        # b1, b2, eps) for a deferred-row backward to apply Adam itself
        # (backward_adam); adam_folded records that it did, this step
        self.adam_fold = None
        self.adam_folded = False

    def keep_gradient_buffer(self, P):
        self.g_theta = torch.zeros((self.m, P), dtype=torch.float32, device=self.x.device)


def theta_gradient(plan, theta):
    """d loss / d theta from the per-chunk rows in plan.partial, or None when
    the plan defers the row sums to adam_rows_partial."""
    if plan.defer_rows:
        return None
    g_theta = plan.g_theta
    if g_theta is None or g_theta.shape != theta.shape:
        g_theta = torch.zeros_like(theta)
    reduce_rows(plan.partial, plan.rows, plan.row_chunk_start, g_theta)
    return g_theta


class FBPINNJets(torch.autograd.Function):
    """ujet[J][n] = jets of the POU-combined subdomain networks at the plan's
    points (FBPINN_model before constraining_fn, trainers.py:126-170, with the
    derivatives of _get_ujs, trainers.py:213-247).  Gradient w.r.t. theta
    ([m][P] packed subdomain rows) for the plan's trained rows only."""

    @staticmethod
    def forward(ctx, theta, plan):
        net, wl = plan.net, plan.wl
        forward(net.as_list(), theta, plan.dom, plan.x, wl.tensor_list(), wl.size_list(), plan.pairbuf)
        ujet = torch.empty((net.J, wl.n_points), dtype=torch.float32, device=theta.device)
        pou_reduce(net.as_list(), wl.tensor_list(), wl.size_list(), plan.pairbuf, plan.wsum, ujet)
        ctx.plan = plan
        ctx.save_for_backward(theta)
        return ujet

    @staticmethod
    def backward(ctx, g_ujet):
        (theta,) = ctx.saved_tensors
        plan = ctx.plan
        net, wl = plan.net, plan.wl
        g_ujet = g_ujet.contiguous()
        pou_backward(net.as_list(), wl.tensor_list(), wl.size_list(), plan.wsum, g_ujet, plan.gseg)
        backward(net.as_list(), theta, plan.dom, plan.x, wl.tensor_list(), wl.size_list(), plan.chunk_train,
                 plan.gseg, plan.partial)
        return theta_gradient(plan, theta), None
