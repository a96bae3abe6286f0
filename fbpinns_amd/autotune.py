"""Launch-shape autotuner for the jet kernels.

fbpx_forward and fbpx_backward each launch one workgroup per bin, a bin being
a contiguous run of chunks (see fbpx_worklist).  Two families of launch shapes:

  ("chunks", C)  one bin per chunk of at most C pairs: many small workgroups,
                 dispatched by the hardware as slots free up;
  ("bins", k)    persistent: k times as many bins as the GPU holds resident
                 workgroups of that kernel (fbpx_occupancy x CU count), each
                 over an equal share of the pairs.  No launch tail, and far
                 fewer chunks, so fewer weight stagings and partial gradient
                 rows.
  ("waves", C)   one-hidden-layer networks: the ("chunks", C) work list with
                 each wave of a workgroup running a chunk of its own.  Where
                 chunks are short (a few 16-pair sub-tiles), a workgroup per
                 chunk leaves most of its waves idle at the barriers.

tune_plan() builds the work list for every candidate, times one forward and
one backward launch on each (HIP events on the current stream, after a warm-up
launch) and returns the work list with the least forward + backward time
among: each candidate shape for both kernels, persistent bins with each
kernel's best multiplier, and the two mixed forms (one kernel on its best
bins, the other on its best chunk size).  Candidates within TIE (2 %) of the
fastest count as equal -- set-up timings of the wave configuration's
candidates move by about that much from box to box -- and among them the
work list with the fewest chunks wins: each chunk is one partial gradient
row that the Adam launch reads back, a cost outside the timed kernels.  A
pair of different shapes is timed on its own work list, since one kernel's
bin edges split the other's chunks and a chunk cap splits the bins' chunks.
Results are cached per (network and jet shape, subdomain count, pair count
to two significant digits), so a scheduler that revisits a shape does not
tune again.  Work lists below MIN_PAIRS take default_spec without timing:
their step is launch-bound and the timings would be noise.

default_spec for a small work list: ("waves", WAVE_CHUNK) for a
one-hidden-layer network with at least WAVE_MIN_CHUNKS_PER_CU chunks (at
FBPX_CHUNK pairs) per CU averaging at most WAVE_MAX_CHUNK_PAIRS pairs (the
short chunks it was measured on), else DEFAULT_SMALL.  Measured on MI355X, batched
1D inverse oscillators, [1, 32, 1], 23 pairs per chunk on average, steps/s of
("chunks", 512) -> ("waves", 32): 90 chunks 27,100 -> 27,400; 355 chunks
25,200 -> 25,000; 719 chunks 22,300 -> 22,900; 1,490 chunks 20,700 ->
22,200; 3,061 chunks 17,800 -> 20,300; 6,078 chunks 12,600 -> 16,000.  With
few chunks a workgroup's waves shorten each chunk's latency chain by sharing
its sub-tiles; with many, the waves a short chunk leaves idle cost more.
"""

import math

import torch

from fbpinns_amd import ops
from fbpinns_amd._lib import FBPX_CHUNK
from fbpinns_amd.util.logger import logger

CANDIDATES = (("chunks", 256), ("chunks", 512), ("bins", 1), ("bins", 2))
WAVE_CHUNK = 32
WAVE_CANDIDATES = (("waves", WAVE_CHUNK),)   # also timed for one-hidden-layer networks
CHUNKS = [c for c in CANDIDATES if c[0] == "chunks"]
BINS = [c for c in CANDIDATES if c[0] == "bins"]
DEFAULT_SMALL = ("chunks", FBPX_CHUNK)
DEFAULT_LARGE = ("bins", 1)
MIN_PAIRS = 200_000
WAVE_MIN_CHUNKS_PER_CU = 2
WAVE_MAX_CHUNK_PAIRS = 64
TIE = 0.02
_cache = {}


def default_spec(n_pairs, nh=2, n_chunks=0, n_cu=1):
    """Launch shape without timing.  n_chunks: chunks of the work list at
    FBPX_CHUNK pairs; n_cu: the GPU's CU count."""
    if n_pairs >= MIN_PAIRS:
        return DEFAULT_LARGE
    if nh == 1 and n_chunks >= WAVE_MIN_CHUNKS_PER_CU * n_cu and n_pairs <= WAVE_MAX_CHUNK_PAIRS * n_chunks:
        return ("waves", WAVE_CHUNK)
    return DEFAULT_SMALL


def candidates(nh):
    return CANDIDATES + (WAVE_CANDIDATES if nh == 1 else ())


def _key(net, n_pairs, m):
    e = max(0, int(math.floor(math.log10(max(n_pairs, 1)))) - 1)
    return tuple(net.as_list()) + (m, round(n_pairs / 10 ** e) * 10 ** e)


MIN_TIMED_MS = 3.0


def _time(fn, reps):
    """Average ms of fn() over at least `reps` launches after a warm-up one,
    and over at least MIN_TIMED_MS of launches for short kernels (set-up
    timings of ~0.1 ms kernels move by several per cent between single
    batches while the clocks ramp)."""
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    fn()
    b.record(s)
    b.synchronize()
    reps = max(reps, min(50, int(MIN_TIMED_MS / max(a.elapsed_time(b), 1e-3)) + 1))
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / reps


def _time_launches(net, theta, dom, x, wplan, reps):
    """(forward ms, backward ms) of one launch each on this work list."""
    wl = wplan.wl
    pairbuf = torch.empty((max(wl.n_pairs, 1), net.J), dtype=torch.float32, device=x.device)
    partial = torch.empty((max(wl.n_chunks, 1), net.P), dtype=torch.float32, device=x.device)
    gseg = torch.zeros((max(wl.n_segs, 1), net.J), dtype=torch.float32, device=x.device)
    args = (net.as_list(), theta, dom, x, wl.tensor_list(), wl.size_list())
    tf = _time(lambda: ops.forward(*args, pairbuf), reps)
    tb = _time(lambda: ops.backward(*args, wplan.chunk_train, gseg, partial), reps)
    return tf, tb


def name(spec):
    if spec[0] == "waves":
        return f"{spec[1]}-pair chunks, one wave each"
    return f"{spec[1]}-pair chunks" if spec[0] == "chunks" else f"{spec[1]}x resident bins"


def tune_plan(build, net, theta, dom, x, m, n_pairs, reps=3, n_chunks=0, n_cu=1):
    """build(fspec, bspec) -> WorkPlan for the current takes.  Returns
    ((fspec, bspec), WorkPlan) of the fastest launch pair.  n_chunks, n_cu:
    as for default_spec, which a work list below MIN_PAIRS takes."""
    if n_pairs < MIN_PAIRS:
        spec = default_spec(n_pairs, net.nh, n_chunks, n_cu)
        return (spec, spec), build(spec, spec)
    key = _key(net, n_pairs, m)
    if key in _cache:
        fs, bs = _cache[key]
        return (fs, bs), build(fs, bs)
    cands = candidates(net.nh)
    tf, tb, nchunks = {}, {}, {}
    for spec in cands:
        plan = build(spec, spec)
        nchunks[(spec, spec)] = plan.wl.n_chunks
        tf[spec], tb[spec] = _time_launches(net, theta, dom, x, plan, reps)

    def measure(f, b):
        plan = build(f, b)
        nchunks[(f, b)] = plan.wl.n_chunks
        return _time_launches(net, theta, dom, x, plan, reps)

    def chunks(pair):   # every compared pair was built above; any other is built here
        if pair not in nchunks:
            nchunks[pair] = build(*pair).wl.n_chunks
        return nchunks[pair]
    (fs, bs), cand = select(tf, tb, measure, chunks=chunks, cands=cands)
    torch.cuda.synchronize()
    _cache[key] = (fs, bs)
    mixed = [k for k in cand if k[0] != k[1]]
    logger.info("autotune launch shape: " + ", ".join(
        f"{name(c)}: fwd {tf[c]:.3f} ms bwd {tb[c]:.3f} ms" for c in cands)
        + "".join(f", fwd {name(f)} + bwd {name(b)}: {cand[(f, b)][0]:.3f} + {cand[(f, b)][1]:.3f} ms"
                  for f, b in mixed)
        + f" -> fwd {name(fs)}, bwd {name(bs)}")
    return (fs, bs), build(fs, bs)


def select(tf, tb, measure, chunks=None, cands=CANDIDATES):
    """Launch pair (forward spec, backward spec) with the least forward +
    backward time among every candidate of `cands` for both kernels and the
    mixed pairs (best bins per kernel; best bins with the other kernel's best
    chunk size).  tf, tb: per-candidate times measured with both kernels on that
    candidate; measure(fspec, bspec) -> (forward ms, backward ms) on the
    mixed pair's own work list.  chunks(pair) -> the chunk count of that
    pair's work list, or None: then the pairs within TIE of the fastest are
    equal and the fewest chunks wins (the fastest on a tie).  Returns (pair, {pair:
    (forward ms, backward ms)})."""
    cf, cb = min(CHUNKS, key=lambda c: tf[c]), min(CHUNKS, key=lambda c: tb[c])
    fbin, bbin = min(BINS, key=lambda c: tf[c]), min(BINS, key=lambda c: tb[c])
    # every uniform pair as measured; a mixed pair on its own work list, since
    # one kernel's bin edges also cut the other kernel's chunks
    cand = {(c, c): (tf[c], tb[c]) for c in cands}
    for pair in ((fbin, bbin), (fbin, cb), (cf, bbin)):
        if pair not in cand:
            cand[pair] = tuple(measure(*pair))
    best = min(cand, key=lambda k: sum(cand[k]))
    if chunks is None:
        return best, cand
    near = [k for k in cand if sum(cand[k]) <= sum(cand[best]) * (1 + TIE)]
    return min(near, key=lambda k: (chunks(k), sum(cand[k]))), cand
