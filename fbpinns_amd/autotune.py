"""Launch-shape autotuner for the jet kernels.

The work list's chunk size (pairs per workgroup of fbpx_forward /
fbpx_backward) trades per-chunk costs (staging the subdomain's weights in
LDS, one gradient row per chunk to write and reduce) against load balance
and the tail of the launch.  tune_chunk() builds the work list for each
candidate size, times one forward and one backward launch on it (HIP events
on the current stream, after a warm-up launch) and returns the fastest.
Results are cached per (network and jet shape, subdomain count, pair count
to two significant digits), so a scheduler that revisits a shape does not
tune again.  Work lists below MIN_PAIRS keep the default size: their step is
launch-bound and the timings would be noise.
"""

import math

import torch

from fbpinns_amd import ops
from fbpinns_amd._lib import FBPX_CHUNK
from fbpinns_amd.util.logger import logger

CANDIDATES = (128, 256, 512, 1024, 2048)
MIN_PAIRS = 200_000
_cache = {}


def _key(net, n_pairs, m):
    e = max(0, int(math.floor(math.log10(max(n_pairs, 1)))) - 1)
    return tuple(net.as_list()) + (m, round(n_pairs / 10 ** e) * 10 ** e)


def _time_launches(net, theta, dom, x, wplan, reps):
    wl = wplan.wl
    pairbuf = torch.empty((max(wl.n_pairs, 1), net.J), dtype=torch.float32, device=x.device)
    partial = torch.empty((max(wl.n_chunks, 1), net.P), dtype=torch.float32, device=x.device)
    gseg = torch.zeros((max(wl.n_segs, 1), net.J), dtype=torch.float32, device=x.device)
    args = (net.as_list(), theta, dom, x, wl.tensor_list(), wl.size_list())

    def once():
        ops.forward(*args, pairbuf)
        ops.backward(*args, wplan.chunk_train, gseg, partial)

    once()
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        once()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / reps


def tune_chunk(build, net, theta, dom, x, m, n_pairs, reps=2):
    """build(chunk) -> WorkPlan for the current takes.  Returns (chunk,
    WorkPlan) with the fastest chunk size among CANDIDATES."""
    if n_pairs < MIN_PAIRS:
        return FBPX_CHUNK, build(FBPX_CHUNK)
    key = _key(net, n_pairs, m)
    if key in _cache:
        c = _cache[key]
        return c, build(c)
    best = None
    times = {}
    for c in CANDIDATES:
        wp = build(c)
        t = _time_launches(net, theta, dom, x, wp, reps)
        times[c] = t
        if best is None or t < best[1]:
            best = (c, t, wp)
    torch.cuda.synchronize()
    _cache[key] = best[0]
    logger.info("autotune chunk size: " + ", ".join(f"{c}: {t:.3f} ms" for c, t in times.items())
                + f" -> {best[0]}")
    return best[0], best[2]
