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
launch) and returns the work list of the faster family: the best chunk size
(forward + backward), or persistent bins with each kernel's best multiplier.
The families are not mixed: another kernel's bin edges would split the chunks
of a per-chunk launch into fragments.  Results are cached per (network and jet
shape, subdomain count, pair count to two significant digits), so a scheduler
that revisits a shape does not tune again.  Work lists below MIN_PAIRS take
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
_cache = {}


def default_spec(n_pairs):
    return DEFAULT_SMALL if n_pairs < MIN_PAIRS else DEFAULT_LARGE


def _key(net, n_pairs, m):
    e = max(0, int(math.floor(math.log10(max(n_pairs, 1)))) - 1)
    return tuple(net.as_list()) + (m, round(n_pairs / 10 ** e) * 10 ** e)


def _time(fn, reps):
    fn()
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
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
    ((fspec, bspec), WorkPlan) of the fastest launch family."""
    if n_pairs < MIN_PAIRS:
        spec = DEFAULT_SMALL
        return (spec, spec), build(spec, spec)
    key = _key(net, n_pairs, m)
    if key in _cache:
        fs, bs = _cache[key]
        return (fs, bs), build(fs, bs)
    tf, tb = {}, {}
    for spec in CANDIDATES:
        tf[spec], tb[spec] = _time_launches(net, theta, dom, x, build(spec, spec), reps)
    torch.cuda.synchronize()
    cs = min(CHUNKS, key=lambda c: tf[c] + tb[c])
    fbin, bbin = min(BINS, key=lambda c: tf[c]), min(BINS, key=lambda c: tb[c])
    fs, bs = (fbin, bbin) if tf[fbin] + tb[bbin] < tf[cs] + tb[cs] else (cs, cs)
    _cache[key] = (fs, bs)
    logger.info("autotune launch shape: " + ", ".join(
        f"{name(c)}: fwd {tf[c]:.3f} ms bwd {tb[c]:.3f} ms" for c in CANDIDATES)
        + f" -> fwd {name(fs)}, bwd {name(bs)}")
    return (fs, bs), build(fs, bs)
