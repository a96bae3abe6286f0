"""ctypes binding of libfbpx.so (the C ABI declared in include/fbpx.h).

The library is built in-tree by fbpinns_amd/csrc/build.py.  There is no
fallback: if the library is missing or a call fails, a FbpxError is raised.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfbpx.so")

FBPX_DOM_STRIDE = 12
FBPX_MAX_XD = 3
FBPX_CHUNK = 512
FBPX_MAX_REQ = 16
FBPX_MAX_TERMS = 8
FBPX_RESIDUAL_BLOCKS = 1024
FBPX_POU_LOSS_RES = 4
FBPX_POU_LOSS_COLS = 8

_c_i32 = ctypes.c_int32
_c_p = ctypes.c_void_p
_c_d = ctypes.c_double


class FbpxError(RuntimeError):
    pass


class fbpx_net(ctypes.Structure):
    _fields_ = [("xd", _c_i32), ("h", _c_i32), ("nh", _c_i32), ("d1mask", _c_i32), ("d2mask", _c_i32)]


class fbpx_worklist(ctypes.Structure):
    _fields_ = [
        ("n_pairs", _c_i32), ("n_chunks", _c_i32), ("n_points", _c_i32), ("n_segs", _c_i32), ("npou", _c_i32),
        ("n_fbins", _c_i32), ("n_bbins", _c_i32),
        ("pair_point", _c_p), ("pair_seg", _c_p), ("chunk_model", _c_p), ("chunk_start", _c_p),
        ("point_seg_start", _c_p), ("seg_pair_start", _c_p), ("pair_slot", _c_p),
        ("fbin_chunk_start", _c_p), ("bbin_chunk_start", _c_p),
        ("fwd_wave_bins", _c_i32), ("bwd_wave_bins", _c_i32),
    ]


class fbpx_takes_out(ctypes.Structure):
    _fields_ = [(k, _c_p) for k in (
        "m_take", "n_take", "p_take", "np_take", "active", "all_ims", "pair_point", "pair_seg", "pair_slot",
        "point_seg_start", "seg_pair_start", "chunk_model", "chunk_start", "chunk_train", "row_chunk_start",
        "fbin_chunk_start", "bbin_chunk_start")]


class fbpx_takes_info(ctypes.Structure):
    _fields_ = [("n_pairs", ctypes.c_int64), ("n_segs", _c_i32), ("n_used", _c_i32), ("n_active", _c_i32),
                ("n_chunks", _c_i32), ("n_fbins", _c_i32), ("n_bbins", _c_i32), ("npou", _c_i32)]


class fbpx_residual(ctypes.Structure):
    _fields_ = [("offset", _c_i32), ("n", _c_i32), ("n_cols", _c_i32), ("n_terms", _c_i32),
                ("weight", ctypes.c_float), ("coef", _c_p), ("fa", _c_i32 * FBPX_MAX_TERMS),
                ("fb", _c_i32 * FBPX_MAX_TERMS), ("jet", _c_i32 * FBPX_POU_LOSS_COLS),
                ("col", _c_p * FBPX_POU_LOSS_COLS), ("col_idx", _c_p * FBPX_POU_LOSS_COLS),
                ("gcol", _c_p * FBPX_POU_LOSS_COLS), ("r", _c_p)]


# name -> (restype, argtypes); mirrors include/fbpx.h
SIGNATURES = {
    "fbpx_inside_count": (_c_i32, [_c_p, _c_p, _c_i32, _c_i32, _c_p, _c_p, _c_i32, _c_p, _c_p]),
    "fbpx_inside_fill": (_c_i32, [_c_p, _c_p, _c_i32, _c_i32, _c_p, _c_p, _c_i32, _c_p, _c_p, _c_p, _c_p]),
    "fbpx_takes_max_chunks": (ctypes.c_int64, [_c_i32, ctypes.c_int64, _c_i32, _c_i32, _c_i32]),
    "fbpx_takes_work_bytes": (ctypes.c_int64, [_c_i32, _c_i32, _c_i32, ctypes.c_int64, _c_i32, _c_i32]),
    "fbpx_takes_count": (_c_i32, [_c_p, _c_p, _c_i32, _c_i32, _c_p, _c_p, _c_i32, _c_p, _c_p, ctypes.c_int64,
                                  ctypes.POINTER(ctypes.c_int64)]),
    "fbpx_build_worklist": (_c_i32, [_c_p, _c_p, _c_i32, _c_i32, _c_p, _c_p, _c_i32, _c_p, _c_p, _c_p,
                                     ctypes.c_int64, _c_i32, _c_i32, _c_i32, ctypes.c_int64, _c_p, ctypes.c_int64,
                                     ctypes.POINTER(fbpx_takes_out), ctypes.POINTER(fbpx_takes_info)]),
    "fbpx_pou_prepare": (_c_i32, [_c_p, ctypes.POINTER(fbpx_net), _c_p, _c_p, ctypes.POINTER(fbpx_worklist),
                                  _c_p, _c_p]),
    "fbpx_forward": (_c_i32, [_c_p, ctypes.POINTER(fbpx_net), _c_p, _c_p, _c_p, ctypes.POINTER(fbpx_worklist),
                              _c_p]),
    "fbpx_pou_reduce": (_c_i32, [_c_p, ctypes.POINTER(fbpx_net), ctypes.POINTER(fbpx_worklist), _c_p, _c_p, _c_p,
                                 _c_p, _c_p]),
    "fbpx_pou_backward": (_c_i32, [_c_p, ctypes.POINTER(fbpx_net), ctypes.POINTER(fbpx_worklist), _c_p, _c_p,
                                   _c_p, _c_p, _c_p]),
    "fbpx_occupancy": (_c_i32, [ctypes.POINTER(fbpx_net), ctypes.POINTER(_c_i32), ctypes.POINTER(_c_i32)]),
    "fbpx_backward": (_c_i32, [_c_p, ctypes.POINTER(fbpx_net), _c_p, _c_p, _c_p, ctypes.POINTER(fbpx_worklist),
                               _c_p, _c_p, _c_p, _c_p]),
    "fbpx_backward_adam": (_c_i32, [_c_p, ctypes.POINTER(fbpx_net), _c_p, _c_p, _c_p, ctypes.POINTER(fbpx_worklist),
                                    _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double]),
    "fbpx_reduce_rows": (_c_i32, [_c_p, _c_i32, _c_p, _c_p, _c_p, _c_i32, _c_p]),
    "fbpx_adam_rows": (_c_i32, [_c_p, _c_p, _c_p, _c_p, _c_i32, _c_p, _c_p, _c_i32, _c_p, _c_p,
                                _c_d, _c_d, _c_d, _c_d]),
    "fbpx_adam_rows_partial": (_c_i32, [_c_p, _c_p, _c_p, _c_p, _c_i32, _c_p, _c_p, _c_p, _c_i32, _c_p, _c_p,
                                        _c_d, _c_d, _c_d, _c_d]),
    "fbpx_adam_flat": (_c_i32, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_i32, _c_p, _c_p, _c_d, _c_d, _c_d, _c_d]),
    "fbpx_count_increment": (_c_i32, [_c_p, _c_p]),
    "fbpx_mean": (_c_i32, [_c_p, _c_p, ctypes.c_int64, _c_p, _c_p]),
    "fbpx_affine_jets": (_c_i32, [_c_p, ctypes.POINTER(fbpx_net), _c_i32, _c_i32, _c_p, _c_p, _c_p, _c_i32,
                                  ctypes.POINTER(_c_i32), ctypes.POINTER(_c_p)]),
    "fbpx_affine_jets_backward": (_c_i32, [_c_p, ctypes.POINTER(fbpx_net), _c_i32, _c_i32, _c_p, _c_i32,
                                           ctypes.POINTER(_c_i32), ctypes.POINTER(_c_p), _c_p]),
    "fbpx_residual_loss": (_c_i32, [_c_p, ctypes.c_int64, _c_i32, ctypes.POINTER(_c_p), ctypes.POINTER(_c_p),
                                    _c_i32, ctypes.POINTER(_c_i32), ctypes.POINTER(_c_i32), _c_p, ctypes.c_float,
                                    _c_p, _c_p, _c_p, _c_i32, _c_p]),
    "fbpx_residual_loss_grad": (_c_i32, [_c_p, ctypes.c_int64, _c_i32, ctypes.POINTER(_c_p), ctypes.POINTER(_c_p),
                                         _c_i32, ctypes.POINTER(_c_i32), ctypes.POINTER(_c_i32), _c_p,
                                         ctypes.c_float, _c_p, _c_p, _c_p, _c_i32, _c_p, ctypes.POINTER(_c_p)]),
    "fbpx_pou_loss": (_c_i32, [_c_p, ctypes.POINTER(fbpx_net), ctypes.POINTER(fbpx_worklist), _c_p, _c_p, _c_p,
                               _c_p, _c_i32, ctypes.POINTER(fbpx_residual), _c_p, _c_p, _c_p, ctypes.c_int64, _c_p,
                               _c_i32, _c_p]),
    "fbpx_residual_losses": (_c_i32, [_c_p, _c_i32, ctypes.POINTER(fbpx_residual), _c_p, ctypes.c_int64, _c_p,
                                      _c_i32, _c_p]),
    "fbpx_residual_loss_backward": (_c_i32, [_c_p, ctypes.c_int64, _c_i32, ctypes.POINTER(_c_p),
                                             ctypes.POINTER(_c_p), _c_i32, ctypes.POINTER(_c_i32),
                                             ctypes.POINTER(_c_i32), _c_p, _c_p, ctypes.c_float, _c_p,
                                             ctypes.POINTER(_c_p), _c_p, _c_p, _c_p]),
    "fbpx_residual_loss_noalloc": (_c_i32, [_c_p, ctypes.c_int64, _c_i32, ctypes.POINTER(_c_p), ctypes.POINTER(_c_p),
                                            _c_i32, ctypes.POINTER(_c_i32), ctypes.POINTER(_c_i32), _c_p,
                                            ctypes.c_float, _c_p, _c_p, _c_i32]),
    "fbpx_segment_sum": (_c_i32, [_c_p, _c_i32, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "fbpx_pou_backward_segment_sum": (_c_i32, [_c_p, ctypes.POINTER(fbpx_net), ctypes.POINTER(fbpx_worklist), _c_p,
                                               _c_p, _c_p, _c_p, _c_p, _c_i32, _c_p, _c_p, _c_p, _c_p]),
    "fbpx_param_count": (_c_i32, [ctypes.POINTER(fbpx_net)]),
    "fbpx_supported": (_c_i32, [ctypes.POINTER(fbpx_net)]),
    "fbpx_last_error": (ctypes.c_char_p, []),
    "fbpx_version": (ctypes.c_char_p, []),
}

_lib = None


def lib():
    """Load libfbpx.so once; raise FbpxError if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FbpxError(f"{LIB_PATH} is not built: run `python fbpinns_amd/csrc/build.py` "
                            "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        msg = lib().fbpx_last_error().decode(errors="replace")
        raise FbpxError(f"{what} failed ({rc}): {msg}")


def param_count(xd, h, nh):
    n = fbpx_net(xd, h, nh, 0, 0)
    return lib().fbpx_param_count(ctypes.byref(n))


def supported(xd, h, nh, d1mask, d2mask):
    n = fbpx_net(xd, h, nh, d1mask, d2mask)
    return bool(lib().fbpx_supported(ctypes.byref(n)))
