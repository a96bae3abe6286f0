"""fbpx_residual_loss_noalloc (work space in a device array of the library)
against fbpx_residual_loss with a caller-allocated work space, both called
through the C ABI on the same inputs: two calls in a row, the second on fewer
points, each with residuals r and loss equal bit for bit (same kernels, same
grid), and the loss within 1e-6 relative of a float64 evaluation."""

import ctypes

import numpy as np
import pytest
import torch

from fbpinns_amd import _lib

pytestmark = pytest.mark.gpu


def _call(name, n, cols, fa, fb, coef, weight, scratch=None):
    L = _lib.lib()
    dev = cols[0].device
    r = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
    loss = torch.full((1,), float("nan"), dtype=torch.float32, device=dev)
    ptrs = (ctypes.c_void_p * len(cols))(*[c.data_ptr() for c in cols])
    a = (ctypes.c_int32 * len(fa))(*fa)
    b = (ctypes.c_int32 * len(fb))(*fb)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if name == "fbpx_residual_loss":
        rc = L.fbpx_residual_loss(stream, n, len(cols), ptrs, None, len(fa), a, b, coef.data_ptr(), weight,
                                  r.data_ptr(), scratch.data_ptr(), loss.data_ptr(), 0, None)
    else:
        rc = L.fbpx_residual_loss_noalloc(stream, n, len(cols), ptrs, None, len(fa), a, b, coef.data_ptr(), weight,
                                          r.data_ptr(), loss.data_ptr(), 0)
    _lib.check(rc, name)
    torch.cuda.synchronize()
    return r.cpu(), loss.cpu()


def test_noalloc_residual_loss_matches_allocating_entry_point():
    g = torch.Generator().manual_seed(11)
    dev = "cuda:0"
    # wave-like residual u_xx + u_yy - u_tt / c^2 (three one-column terms) and
    # a constant term, over two sizes: many workgroups, then few
    fa, fb = [0, 1, 2, -1], [-1, -1, 3, -1]
    coef = torch.tensor([1.0, 1.0, -1.0, 0.25], dtype=torch.float32, device=dev)
    weight = 0.7
    for n in (1_000_003, 3_001):
        cols = [torch.randn(n, generator=g).to(dev) for _ in range(4)]
        print(f"[noalloc] n={n}", flush=True)
        scratch = torch.empty(_lib.FBPX_RESIDUAL_BLOCKS, dtype=torch.float32, device=dev)
        r0, l0 = _call("fbpx_residual_loss", n, cols, fa, fb, coef, weight, scratch)
        r1, l1 = _call("fbpx_residual_loss_noalloc", n, cols, fa, fb, coef, weight)
        c = [x.double().cpu() for x in cols]
        cf = coef.double().cpu()
        rr = cf[0] * c[0] + cf[1] * c[1] + cf[2] * c[2] * c[3] + cf[3]
        lref = weight / n * (rr ** 2).sum().item()
        print(f"  loss allocating {l0.item():.9e} noalloc {l1.item():.9e} float64 {lref:.9e}", flush=True)
        assert torch.equal(r0, r1)
        assert torch.equal(l0, l1)
        assert abs(l1.item() - lref) <= 1e-6 * abs(lref)
        assert np.allclose(r1.double().numpy(), rr.numpy(), rtol=1e-6, atol=1e-6)


def test_noalloc_rejects_bad_sizes():
    L = _lib.lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    x = torch.zeros(4, device="cuda:0")
    ptrs = (ctypes.c_void_p * 1)(x.data_ptr())
    a = (ctypes.c_int32 * 9)(*([0] * 9))
    for n, nt in ((0, 1), (4, 0), (4, _lib.FBPX_MAX_TERMS + 1)):
        rc = L.fbpx_residual_loss_noalloc(stream, n, 1, ptrs, None, nt, a, a, x.data_ptr(), 1.0, x.data_ptr(),
                                          x.data_ptr(), 0)
        assert rc == -1, (n, nt, rc)
        assert L.fbpx_last_error().decode()
