"""Take builder and work-list construction.

device_plan: the takes of get_inputs (trainers.py:332-419) and the kernels'
work list (include/fbpx.h: fbpx_worklist) built on the GPU in one pass by
fbpx_takes_count / fbpx_build_worklist; the train step and inference use it.

get_inputs / get_x_batch / split_takes keep the reference's host-side form
and output format (trainers.py:332-419, 482-507, 544-571), with the O(n m)
inside tests on the GPU (fbpx_inside_count / fbpx_inside_fill) and the integer
bookkeeping in numpy: the plain per-op form (plain.py) runs the reference's
method with them, and get_x_batch selects the training points of a
scheduler state for both.
"""

import numpy as np
import torch

from fbpinns_amd import ops
from fbpinns_amd.ops import Worklist


def _tree_cut(d, idx):
    """Cut the leading (subdomain) axis of every leaf under a 'subdomain' key."""
    def cut(v):
        if isinstance(v, dict):
            return {k: cut(x) for k, x in v.items()}
        if isinstance(v, (list, tuple)):
            return type(v)(cut(x) for x in v)
        if isinstance(v, torch.Tensor):
            return v[torch.as_tensor(idx, dtype=torch.long, device=v.device)]
        if isinstance(v, np.ndarray):
            return v[idx]
        return v
    return {cl: {k: (cut(v) if k == "subdomain" else v) for k, v in d[cl].items()} for cl in d}


def inputs_from_inside(n_take, m_take, training_ims, active, pou_all):
    """Integer part of get_inputs (trainers.py:347-391) given the inside
    takes: active remap, all_ims = [active; fixed], m_take re-index and the
    (point, pou) unique.  Returns takes, all_ims, active, active_ims, fixed_ims."""
    m = len(pou_all)
    active = np.array(active, dtype=np.int64).copy()
    assert np.isin(active, [0, 1, 2]).all()
    assert active.shape == (m,)
    active[active == 0] = 1
    mask = np.zeros_like(active)
    mask[training_ims] = 1
    active = active * mask
    ims_ = np.arange(m)
    active_ims = ims_[active == 1]
    fixed_ims = ims_[active == 2]
    all_ims = np.concatenate([active_ims, fixed_ims])
    inv = np.zeros(m, dtype=np.int64)
    inv[all_ims] = np.arange(len(all_ims))
    m_take = inv[m_take]
    pou_all = np.asarray(pou_all).astype(np.int64).reshape(-1)
    pous = pou_all[all_ims]
    # jnp.unique over (point, pou) rows (lexicographic order) with inverse,
    # through the equivalent 1-D key point * npou_max + pou
    pt = pous[m_take]
    npou_max = int(pou_all.max()) + 1 if len(pou_all) else 1
    ukey, p_take = np.unique(n_take * npou_max + pt, return_inverse=True)
    p_take = p_take.reshape(-1)
    np_take = ukey // npou_max
    npou = len(np.unique(pou_all))
    return (m_take, n_take, p_take, np_take, npou), all_ims, active, active_ims, fixed_ims


def get_inputs(x_batch, active, all_params, decomposition):
    """trainers.py:332-419.  x_batch: (n, xd) float32 GPU tensor.
    Returns takes = (m_take, n_take, p_take, np_take, npou) as int64 numpy
    arrays (m_take indexes all_ims), all_ims, and
    (active, cut_active, cut_fixed, cut_all, merge_active)."""
    n_take, m_take, training_ims = decomposition.inside_points(all_params, x_batch)
    pou_all = np.asarray(all_params["static"]["decomposition"]["subdomain"]["pou"].cpu()).reshape(-1)
    takes, all_ims, active, active_ims, fixed_ims = inputs_from_inside(n_take, m_take, training_ims, active,
                                                                      pou_all)

    def cut_active(d):
        return _tree_cut(d, active_ims)

    def cut_fixed(d):
        return _tree_cut(d, fixed_ims)

    def cut_all(d):
        return _tree_cut(d, all_ims)

    def merge_active(da, d):
        for cl in d:
            for k in d[cl]:
                if k == "subdomain":
                    d[cl][k] = _merge(da[cl][k], d[cl][k], active_ims)
                else:
                    d[cl][k] = da[cl][k]
        return d

    return takes, all_ims, (active, cut_active, cut_fixed, cut_all, merge_active)


def _merge(va, v, idx):
    if isinstance(v, dict):
        return {k: _merge(va[k], v[k], idx) for k in v}
    if isinstance(v, (list, tuple)):
        return type(v)(_merge(a, b, idx) for a, b in zip(va, v))
    out = v.clone()
    out[torch.as_tensor(idx, dtype=torch.long, device=v.device)] = va
    return out


def get_x_batch(active, all_params, x_batch_global, constraint_offsets, constraint_sizes, decomposition):
    """trainers.py:482-507 index bookkeeping: training_ips (points inside >= 1
    strictly active model), constraint_fs (n_train, c), constraint_ips and d."""
    active = np.asarray(active)
    ims = np.arange(all_params["static"]["decomposition"]["m"])[active == 1]
    training_ips, d = decomposition.inside_models(all_params, x_batch_global, ims)
    n_glob = x_batch_global.shape[0]
    fs_global = np.zeros((n_glob, len(constraint_sizes)), dtype=bool)
    for ic, (o, s) in enumerate(zip(constraint_offsets, constraint_sizes)):
        fs_global[o:o + s, ic] = True
    constraint_fs = fs_global[training_ips]
    ix_ = np.arange(len(training_ips))
    constraint_ips = [ix_[f] for f in constraint_fs.T]
    return training_ips, constraint_fs, constraint_ips, d


def split_takes(takes, constraint_fs, constraint_ips, n_train):
    """trainers.py:544-571: per-constraint re-indexed takes."""
    m_take, n_take, p_take, np_take, npou = takes
    takess = []
    iu_ = np.arange(np_take.shape[0])
    for f, ips in zip(constraint_fs.T, constraint_ips):
        f2 = f[np_take]
        ius = iu_[f2]
        f3 = f[n_take]
        f4 = f2[p_take]
        inv = np.zeros(n_train, dtype=np.int64)
        inv[ips] = np.arange(len(ips))
        inv2 = np.zeros(np_take.shape[0], dtype=np.int64)
        inv2[ius] = np.arange(len(ius))
        takess.append((m_take[f3], inv[n_take[f3]], inv2[p_take[f4]], inv[np_take[f2]], npou))
    return takess


def spatial_order(x):
    """Permutation of the rows of x (n, xd) along a Morton (Z-order) curve over
    their bounding box (21 bits per axis for xd = 1 to 3, ties in input order).

    In this order the points of one subdomain form a few contiguous runs, so
    the pair kernels' gathers of points and of upstream jets, and the forward's
    scattered writes of pair jets, share cache lines instead of touching one
    line per pair."""
    x = np.asarray(x, dtype=np.float64)
    n, xd = x.shape
    if n < 2:
        return np.arange(n)
    bits = min(21, 63 // max(xd, 1))
    lo, hi = x.min(0), x.max(0)
    span = np.where(hi > lo, hi - lo, 1.0)
    q = np.floor((x - lo) / span * ((1 << bits) - 1)).astype(np.uint64)
    code = np.zeros(n, dtype=np.uint64)
    for b in range(bits):
        for i in range(xd):
            code |= ((q[:, i] >> np.uint64(b)) & np.uint64(1)) << np.uint64(b * xd + i)
    return np.argsort(code, kind="stable")


class WorkPlan:
    """Work list plus the row bookkeeping of one set of takes."""

    def __init__(self, wl, chunk_train, rows, row_chunk_start, n_active, all_ims):
        self.wl, self.chunk_train, self.rows, self.row_chunk_start = wl, chunk_train, rows, row_chunk_start
        self.n_active, self.all_ims = n_active, all_ims


class Takes:
    """The reference-format takes (m_take, n_take, p_take, np_take, npou) of
    a device-built work list, kept on the GPU; indexing copies the four
    arrays to int64 numpy once."""

    def __init__(self, t, npou):
        self.t, self.npou, self._np = t, int(npou), None

    def numpy(self):
        if self._np is None:
            self._np = tuple(self.t[k].cpu().numpy().astype(np.int64)
                             for k in ("m_take", "n_take", "p_take", "np_take")) + (self.npou,)
        return self._np

    def __getitem__(self, i):
        return self.numpy()[i]

    def __len__(self):
        return 5

    def __iter__(self):
        return iter(self.numpy())


_WL_KEYS = ("pair_point", "pair_seg", "chunk_model", "chunk_start", "point_seg_start", "seg_pair_start",
            "pair_slot", "fbin_chunk_start", "bbin_chunk_start")


def device_plan(x, active, all_params, chunk=None, fbins=None, bbins=None):
    """get_inputs (trainers.py:332-419) and the work list of its takes, built
    on the GPU (fbpx_takes_count + fbpx_build_worklist).

    Tile order groups the pairs by position in all_ims = [active; fixed],
    ascending point inside a subdomain, so the trained pairs are a prefix.
    Chunks (runs of one subdomain's pairs, one partial gradient row each) end
    at every subdomain boundary, every `chunk` pairs inside a subdomain when
    `chunk` is set, and at the bin edges.  fbins / bbins: the forward /
    backward launch as that many workgroups, each over an equal share of the
    pairs (the backward over the trained prefix only), a contiguous run of
    chunks per workgroup; None = one workgroup per chunk.  Segments are the
    unique (point, pou) rows; pair_slot gives each tile-order pair its slot in
    segment order (a segment's pairs are contiguous there, in the reference's
    pair order).

    x: (n, xd) float32 GPU tensor.  Returns (WorkPlan, Takes, all_ims,
    active) with all_ims and the remapped active vector as int64 numpy."""
    dec = all_params["static"]["decomposition"]
    p = dec["subdomain"]["params"]
    pou = dec["subdomain"]["pou"].reshape(-1).round().to(torch.int32)
    t, info = ops.device_worklist(x, p[0], p[1], np.asarray(active, dtype=np.int32), pou, int(chunk or 0),
                                  int(fbins or 0), int(bbins or 0))
    sizes = {"n_pairs": info["n_pairs"], "n_chunks": info["n_chunks"], "n_points": x.shape[0],
             "n_segs": info["n_segs"], "npou": info["npou"], "n_fbins": info["n_fbins"], "n_bbins": info["n_bbins"]}
    wl = Worklist({k: t[k] for k in _WL_KEYS}, sizes)
    all_ims = t["all_ims"].cpu().numpy().astype(np.int64)
    active_ = t["active"].cpu().numpy().astype(np.int64)
    n_active = info["n_active"]
    plan = WorkPlan(wl, t["chunk_train"], t["all_ims"][:n_active], t["row_chunk_start"], n_active, all_ims)
    return plan, Takes(t, info["npou"]), all_ims, active_
