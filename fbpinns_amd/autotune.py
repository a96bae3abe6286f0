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
tune again.  Work lists below MIN_PAIRS take
DEFAULT_SMALL without timing: their step is launch-bound and the timings would
be noise.
"""

import math

import torch

from fbpinns_amd import ops
from fbpinns_amd._lib import FBPX_CHUNK
from fbpinns_amd.util.logger import logger

CANDIDATES = (("chunks", 256), ("chunks", 512), ("bins", 1), ("bins", 2))
CHUNKS = [c for c in CANDIDATES if c[0] == "chunks"]
BINS = [c for c in CANDIDATES if c[0] == "bins"]
DEFAULT_SMALL = ("chunks", FBPX_CHUNK)
DEFAULT_LARGE = ("bins", 1)
MIN_PAIRS = 200_000
TIE = 0.02
_cache = {}


def default_spec(n_pairs):
    return DEFAULT_SMALL if n_pairs < MIN_PAIRS else DEFAULT_LARGE


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
    return f"{spec[1]}-pair chunks" if spec[0] == "chunks" else f"{spec[1]}x resident bins"


def tune_plan(build, net, theta, dom, x, m, n_pairs, reps=3):
    """build(fspec, bspec) -> WorkPlan for the current takes.  Returns
    ((fspec, bspec), WorkPlan) of the fastest launch pair."""
    if n_pairs < MIN_PAIRS:
        spec = DEFAULT_SMALL
        return (spec, spec), build(spec, spec)
    key = _key(net, n_pairs, m)
    if key in _cache:
        fs, bs = _cache[key]
        return (fs, bs), build(fs, bs)
    tf, tb, nchunks = {}, {}, {}
    for spec in CANDIDATES:
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
    (fs, bs), cand = select(tf, tb, measure, chunks=chunks)
    torch.cuda.synchronize()
    _cache[key] = (fs, bs)
    mixed = [k for k in cand if k[0] != k[1]]
    logger.info("autotune launch shape: " + ", ".join(
        f"{name(c)}: fwd {tf[c]:.3f} ms bwd {tb[c]:.3f} ms" for c in CANDIDATES)
        + "".join(f", fwd {name(f)} + bwd {name(b)}: {cand[(f, b)][0]:.3f} + {cand[(f, b)][1]:.3f} ms"
                  for f, b in mixed)
        + f" -> fwd {name(fs)}, bwd {name(bs)}")
    return (fs, bs), build(fs, bs)


def select(tf, tb, measure, chunks=None):
    """Launch pair (forward spec, backward spec) with the least forward +
    backward time among every candidate for both kernels and the mixed pairs
    (best bins per kernel; best bins with the other kernel's best chunk
    size).  tf, tb: per-candidate times measured with both kernels on that
    candidate; measure(fspec, bspec) -> (forward ms, backward ms) on the
    mixed pair's own work list.  chunks(pair) -> the chunk count of that
    pair's work list, or None: then the pairs within TIE of the fastest are
    equal and the fewest chunks wins (the fastest on a tie).  Returns (pair, {pair:
    (forward ms, backward ms)})."""
    cf, cb = min(CHUNKS, key=lambda c: tf[c]), min(CHUNKS, key=lambda c: tb[c])
    fbin, bbin = min(BINS, key=lambda c: tf[c]), min(BINS, key=lambda c: tb[c])
    # every uniform pair as measured; a mixed pair on its own work list, since
    # one kernel's bin edges also cut the other kernel's chunks
    cand = {(c, c): (tf[c], tb[c]) for c in CANDIDATES}
    for pair in ((fbin, bbin), (fbin, cb), (cf, bbin)):
        if pair not in cand:
            cand[pair] = tuple(measure(*pair))
    best = min(cand, key=lambda k: sum(cand[k]))
    if chunks is None:
        return best, cand
    near = [k for k in cand if sum(cand[k]) <= sum(cand[best]) * (1 + TIE)]
    return min(near, key=lambda k: (chunks(k), sum(cand[k]))), cand
