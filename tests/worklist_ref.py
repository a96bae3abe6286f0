"""The work-list layout restated on the host in numpy, for the tests: the
device builder (fbpx_build_worklist, fbpinns_amd.takes.device_plan) must
produce these arrays exactly from the reference-format takes."""

import numpy as np
import torch

from fbpinns_amd.ops import Worklist
from fbpinns_amd.takes import WorkPlan


def build_worklist(takes, all_ims, n_active, n_points, device, chunk=512, fbins=None, bbins=None):
    """Kernel work list from the reference-format takes.

    Tile order groups the pairs by position in all_ims = [active; fixed]
    (stable, so ascending point order inside a subdomain), so the trained
    pairs are a prefix.  Chunks (runs of one subdomain's pairs, one partial
    gradient row each) end at every subdomain boundary, every `chunk` pairs
    inside a subdomain when `chunk` is set, and at the bin edges.  fbins /
    bbins: the forward / backward launch as that many workgroups, each over
    an equal share of the pairs (the backward over the trained prefix only),
    a contiguous run of chunks per workgroup; None = one workgroup per chunk.
    Segments are the unique (point, pou) rows; pair_slot gives each
    tile-order pair its slot in segment order (a segment's pairs are
    contiguous there, in the reference's pair order)."""
    m_take, n_take, p_take, np_take, npou = takes
    s = len(m_take)
    n_all = len(all_ims)
    order = np.argsort(m_take, kind="stable")
    counts = np.bincount(m_take, minlength=n_all)
    model_start = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    s_train = int(model_start[n_active])
    cuts = [model_start]
    if chunk is not None:
        chunk = int(chunk)
        nch = (counts + chunk - 1) // chunk
        pos = np.repeat(np.arange(n_all), nch)
        within = np.arange(len(pos)) - np.concatenate([[0], np.cumsum(nch)])[:-1][pos]
        cuts.append(model_start[pos] + within * chunk)
    fedges = bedges = None
    if fbins is not None:
        fedges = np.round(np.linspace(0, s, int(fbins) + 1)).astype(np.int64)
        cuts.append(fedges)
    if bbins is not None:
        bedges = np.round(np.linspace(0, s_train, int(bbins) + 1)).astype(np.int64)
        cuts.append(bedges)
    bounds = np.unique(np.concatenate(cuts))
    bounds = bounds[(bounds >= 0) & (bounds <= s)]
    if len(bounds) == 0 or bounds[-1] != s:
        bounds = np.append(bounds, s)
    chunk_start = bounds.astype(np.int32)                            # [n_chunks + 1], first 0, last s
    n_chunks = len(chunk_start) - 1
    # the subdomain position holding each chunk (empty positions repeat a start; take the last)
    chunk_pos = np.searchsorted(model_start, chunk_start[:-1], side="right") - 1
    chunk_model = np.asarray(all_ims, dtype=np.int64)[chunk_pos].astype(np.int32)
    chunk_train = (chunk_pos < n_active).astype(np.uint8)
    row_chunk_start = np.searchsorted(chunk_pos, np.arange(n_active + 1), side="left").astype(np.int32)
    rows = np.asarray(all_ims[:n_active], dtype=np.int32)
    one_per_chunk = np.arange(n_chunks + 1, dtype=np.int32)
    fbin_chunk_start = one_per_chunk if fedges is None else np.searchsorted(chunk_start, fedges).astype(np.int32)
    bbin_chunk_start = one_per_chunk if bedges is None else np.searchsorted(chunk_start, bedges).astype(np.int32)

    tile_pos = np.empty(s, dtype=np.int64)
    tile_pos[order] = np.arange(s)
    n_segs = len(np_take)
    seg_order = np.argsort(p_take, kind="stable")
    pair_slot = np.empty(s, dtype=np.int32)
    pair_slot[tile_pos[seg_order]] = np.arange(s, dtype=np.int32)
    seg_pair_start = np.concatenate([[0], np.cumsum(np.bincount(p_take, minlength=n_segs))]).astype(np.int32)
    point_seg_start = np.searchsorted(np_take, np.arange(n_points + 1), side="left").astype(np.int32)

    dev = lambda a, dt=torch.int32: torch.as_tensor(np.ascontiguousarray(a)).to(device=device, dtype=dt)
    tensors = {
        "pair_point": dev(n_take[order]),
        "pair_seg": dev(p_take[order]),
        "chunk_model": dev(chunk_model),
        "chunk_start": dev(chunk_start),
        "point_seg_start": dev(point_seg_start),
        "seg_pair_start": dev(seg_pair_start),
        "pair_slot": dev(pair_slot),
        "fbin_chunk_start": dev(fbin_chunk_start),
        "bbin_chunk_start": dev(bbin_chunk_start),
    }
    sizes = {"n_pairs": s, "n_chunks": n_chunks, "n_points": n_points, "n_segs": n_segs, "npou": npou,
             "n_fbins": len(fbin_chunk_start) - 1, "n_bbins": len(bbin_chunk_start) - 1}
    wl = Worklist(tensors, sizes)
    return WorkPlan(wl, dev(chunk_train, torch.uint8), dev(rows), dev(row_chunk_start), n_active,
                    np.asarray(all_ims))
