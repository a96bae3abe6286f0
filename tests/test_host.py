"""CPU checks of the product's host logic against the oracle, the C-ABI
library's exports, the loss epilogue's two paths, and the API surface
(schedulers, constants, domains, decompositions)."""

import ctypes
import os
import re

import numpy as np
import pytest
import torch

import setups
import worklist_ref
from fbpinns_amd import _lib, epilogue, takes as ptakes
from fbpinns_amd.constants import Constants, get_subdomain_ws
from fbpinns_amd.decompositions import RectangularDecompositionND
from fbpinns_amd.domains import RectangularDomainND
from fbpinns_amd.engine import pick_schema
from fbpinns_amd.ops import FbpxError, NetSpec
from fbpinns_amd import problems as P
from fbpinns_amd import schedulers as S
from fbpinns_amd.util import rng
from oracle import model as om
from oracle import takes as ot

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ------------------------------------------------------------------ takes

@pytest.mark.parametrize("seed,ml", [(0, False), (1, False), (2, True)])
def test_inputs_from_inside_matches_oracle(seed, ml):
    r = np.random.default_rng(seed)
    _, dec = setups.decomposition_params((7, 5), (0.0, 0.0), (1.0, 1.0), 1.9, multilevel=ml)
    lo, hi = dec["subdomain"]["params"][0].numpy(), dec["subdomain"]["params"][1].numpy()
    pou = dec["subdomain"]["pou"].numpy().reshape(-1)
    m = lo.shape[0]
    x = r.uniform(0, 1, (500, 2)).astype(np.float32)
    active = r.integers(0, 3, m)
    n_take, m_take, ims = ot.inside_points(x, lo, hi)
    got = ptakes.inputs_from_inside(n_take, m_take, ims, active, pou)
    want = ot.get_inputs(x, active, lo, hi, pou)
    for a, b in zip(got[0][:4], want[0][:4]):
        assert np.array_equal(a, b)
    assert got[0][4] == want[0][4]
    for a, b in zip(got[1:], want[1:]):
        assert np.array_equal(a, b)


def test_split_takes_matches_oracle():
    r = np.random.default_rng(3)
    _, dec = setups.decomposition_params((6,), (0.0,), (1.0,), 2.2)
    lo, hi = dec["subdomain"]["params"][0].numpy(), dec["subdomain"]["params"][1].numpy()
    xg = r.uniform(0, 1, (90, 1)).astype(np.float32)
    offs, sizes = [0, 50, 70], [50, 20, 20]
    tips, local, cfs, cips = ot.get_x_batch(np.ones(6, int), xg, offs, sizes, lo, hi)
    takes, *_ = ot.get_inputs(xg[tips], np.ones(6, int), lo, hi, np.zeros(6))
    a = ptakes.split_takes(takes, cfs, cips, len(tips))
    b = ot.split_takes(takes, cfs, cips, len(tips))
    for ta, tb in zip(a, b):
        for u, v in zip(ta[:4], tb[:4]):
            assert np.array_equal(u, v)


def test_build_worklist_invariants():
    r = np.random.default_rng(5)
    _, dec = setups.decomposition_params((9, 4), (0.0, 0.0), (1.0, 1.0), 2.5, multilevel=True)
    lo, hi = dec["subdomain"]["params"][0].numpy(), dec["subdomain"]["params"][1].numpy()
    pou = dec["subdomain"]["pou"].numpy().reshape(-1)
    m = lo.shape[0]
    x = r.uniform(0, 1, (3000, 2)).astype(np.float32)
    active = np.ones(m, int)
    active[::4] = 2
    takes, all_ims, act, a_ims, f_ims = ot.get_inputs(x, active, lo, hi, pou)
    plan = worklist_ref.build_worklist(takes, all_ims, len(a_ims), x.shape[0], "cpu", chunk=64)
    t = {k: v.numpy() for k, v in plan.wl.t.items()}
    m_take, n_take, p_take, np_take, npou = takes
    cs = t["chunk_start"]
    assert cs[0] == 0 and cs[-1] == len(m_take) and np.all(np.diff(cs) > 0) and np.all(np.diff(cs) <= 64)
    # every chunk belongs to one subdomain and lists its pairs in ascending point order
    glob = all_ims[m_take]
    pair_model = np.repeat(t["chunk_model"], np.diff(cs))
    order = np.argsort(m_take, kind="stable")
    assert np.array_equal(pair_model, glob[order])
    assert np.array_equal(t["pair_point"], n_take[order])
    # trained chunks come first; row ranges cover exactly the active subdomains' chunks
    ct = plan.chunk_train.numpy().astype(bool)
    assert np.all(ct[:ct.sum()]) and not np.any(ct[ct.sum():])
    rcs = plan.row_chunk_start.numpy()
    assert np.array_equal(plan.rows.numpy(), a_ims)
    for k, row in enumerate(a_ims):
        assert set(t["chunk_model"][rcs[k]:rcs[k + 1]]) == {row}
    # pair_slot is a permutation; the slots of a segment's pairs are its range
    slot = t["pair_slot"]
    assert np.array_equal(np.sort(slot), np.arange(len(m_take)))
    sp = np.empty(len(m_take), np.int64)
    sp[slot] = np.arange(len(m_take))          # slot -> tile-order pair
    inv = np.empty(len(m_take), np.int64)
    inv[order] = np.arange(len(m_take))
    ss = t["seg_pair_start"]
    for sgi in r.integers(0, len(np_take), 40):
        ids = sp[ss[sgi]:ss[sgi + 1]]
        assert np.all(t["pair_seg"][ids] == sgi)
    pss = t["point_seg_start"]
    assert pss[-1] == len(np_take) and np.all(np.diff(pss) >= 1)


def test_build_worklist_bins():
    """Persistent launch bins: the forward's bins split all pairs, the
    backward's the trained prefix, each into equal shares (+-1 pair) of whole
    chunks; chunk boundaries fall on subdomain boundaries and bin edges (and
    every `chunk` pairs when a cap is given)."""
    r = np.random.default_rng(6)
    _, dec = setups.decomposition_params((9, 4), (0.0, 0.0), (1.0, 1.0), 2.5, multilevel=False)
    lo, hi = dec["subdomain"]["params"][0].numpy(), dec["subdomain"]["params"][1].numpy()
    pou = dec["subdomain"]["pou"].numpy().reshape(-1)
    m = lo.shape[0]
    x = r.uniform(0, 1, (2500, 2)).astype(np.float32)
    active = np.ones(m, int)
    active[::3] = 2
    takes, all_ims, act, a_ims, f_ims = ot.get_inputs(x, active, lo, hi, pou)
    s = len(takes[0])
    counts = np.bincount(takes[0], minlength=len(all_ims))
    s_train = int(counts[:len(a_ims)].sum())
    model_start = np.concatenate([[0], np.cumsum(counts)])
    for cap, nf, nb in ((None, 7, 5), (None, 1, 1), (64, 13, 3), (None, 300, 200)):
        plan = worklist_ref.build_worklist(takes, all_ims, len(a_ims), x.shape[0], "cpu", chunk=cap, fbins=nf, bbins=nb)
        t = {k: v.numpy() for k, v in plan.wl.t.items()}
        cs, fb, bb = t["chunk_start"], t["fbin_chunk_start"], t["bbin_chunk_start"]
        n_chunks = plan.wl.n_chunks
        assert plan.wl.n_fbins == nf and plan.wl.n_bbins == nb
        assert cs[0] == 0 and cs[-1] == s and np.all(np.diff(cs) > 0) and len(cs) == n_chunks + 1
        if cap is not None:
            assert np.all(np.diff(cs) <= cap)
        assert set(model_start.tolist()) <= set(cs.tolist())
        glob = all_ims[takes[0]]
        order = np.argsort(takes[0], kind="stable")
        assert np.array_equal(np.repeat(t["chunk_model"], np.diff(cs)), glob[order])
        # forward bins: all chunks, equal pair shares
        assert fb[0] == 0 and fb[-1] == n_chunks and np.all(np.diff(fb) >= 0)
        assert np.all(np.abs(np.diff(cs[fb]) - s / nf) <= 1)
        # backward bins: exactly the trained chunks, equal shares of the trained pairs
        n_train_chunks = int(plan.chunk_train.numpy().sum())
        assert cs[n_train_chunks] == s_train
        assert bb[0] == 0 and bb[-1] == n_train_chunks and np.all(np.diff(bb) >= 0)
        assert np.all(np.abs(np.diff(cs[bb]) - s_train / nb) <= 1)
        rcs = plan.row_chunk_start.numpy()
        for k, row in enumerate(a_ims):
            assert set(t["chunk_model"][rcs[k]:rcs[k + 1]]) <= {row}
            assert cs[rcs[k + 1]] - cs[rcs[k]] == counts[k]
    # no bins: one workgroup per chunk for both kernels
    plan = worklist_ref.build_worklist(takes, all_ims, len(a_ims), x.shape[0], "cpu", chunk=64)
    n_chunks = plan.wl.n_chunks
    assert np.array_equal(plan.wl.t["fbin_chunk_start"].numpy(), np.arange(n_chunks + 1))
    assert np.array_equal(plan.wl.t["bbin_chunk_start"].numpy(), np.arange(n_chunks + 1))


def test_spatial_order():
    """Z-order permutation of training points: 1-D is the stable sort (ties
    keep input order); a 4 x 4 grid given in shuffled order comes out in the
    Z-curve order listed by hand (axis 0 the low bit of each level); for
    uniform 3-D points each subdomain's points form a few long runs instead of
    one run per point."""
    r = np.random.default_rng(3)
    x1 = r.uniform(0, 1, (500, 1))
    x1[100:110] = x1[5]
    assert np.array_equal(ptakes.spatial_order(x1), np.argsort(x1[:, 0], kind="stable"))
    z = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 0), (3, 0), (2, 1), (3, 1),
         (0, 2), (1, 2), (0, 3), (1, 3), (2, 2), (3, 2), (2, 3), (3, 3)]
    shuffle = r.permutation(16)
    g = np.array(z, dtype=np.float64)[shuffle] * 0.25 + 0.1
    out = [tuple(int(round(v)) for v in (g[i] - 0.1) / 0.25) for i in ptakes.spatial_order(g)]
    assert out == z
    assert ptakes.spatial_order(np.zeros((1, 3))).tolist() == [0]
    assert ptakes.spatial_order(np.zeros((0, 2))).tolist() == []
    x = r.uniform(0, 1, (20000, 3))
    order = ptakes.spatial_order(x)
    assert np.array_equal(np.sort(order), np.arange(len(x)))
    _, dec = setups.decomposition_params((6, 6, 3), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), 1.9)
    lo, hi = dec["subdomain"]["params"][0].numpy(), dec["subdomain"]["params"][1].numpy()

    def runs_per_point(o):
        xs = x[o]
        n_runs = n_pts = 0
        for j in range(lo.shape[0]):
            idx = np.nonzero(np.all((xs >= lo[j]) & (xs <= hi[j]), 1))[0]
            n_runs += 1 + int((np.diff(idx) > 1).sum())
            n_pts += len(idx)
        return n_runs / n_pts
    assert runs_per_point(order) < 0.2 < 0.8 < runs_per_point(np.arange(len(x)))


def test_autotune_select():
    """The autotuner's choice: the least forward + backward time among every
    candidate for both kernels and the mixed pairs (per-kernel best bins, and
    best bins with the other kernel's best chunk size), each mixed pair timed
    on its own work list, not assembled from the uniform timings."""
    from fbpinns_amd import autotune as A
    c256, c512, b1, b2 = ("chunks", 256), ("chunks", 512), ("bins", 1), ("bins", 2)
    # the poisson-like case: forward fastest on bins, backward on 256-pair chunks
    tf = {c256: 0.164, c512: 0.122, b1: 0.120, b2: 0.124}
    tb = {c256: 0.313, c512: 0.335, b1: 0.331, b2: 0.352}
    asked = []

    def measure(f, b):
        asked.append((f, b))
        return {(b1, c256): (0.121, 0.314), (c512, b1): (0.123, 0.333)}[(f, b)]
    pair, cand = A.select(tf, tb, measure)
    assert sorted(asked) == sorted([(b1, c256), (c512, b1)])
    assert pair == (b1, c256) and cand[pair] == (0.121, 0.314)
    assert cand[(c512, c512)] == (0.122, 0.335) and cand[(b1, b1)] == (0.120, 0.331)
    # a mixed form that measures slower than its parts loses
    pair, _ = A.select(tf, tb, lambda f, b: (0.2, 0.4))
    assert pair == (b1, b1)
    # the wave-like case: forward fastest on 2x bins, backward on 1x bins.  The
    # mixed pair is timed on its own work list (the forward's bin edges cut the
    # backward's chunks) and loses to uniform 1x bins when that slows its
    # backward; it wins when its own work list keeps both kernels fast
    tf = {c256: 2.95, c512: 2.94, b1: 2.90, b2: 2.86}
    tb = {c256: 7.88, c512: 7.81, b1: 7.65, b2: 7.70}
    mixed = {(b2, b1): (2.86, 7.82), (b2, c512): (2.87, 7.90), (c512, b1): (2.95, 7.66)}
    asked = []

    def measure(f, b):
        asked.append((f, b))
        return mixed[(f, b)]
    pair, cand = A.select(tf, tb, measure)
    assert sorted(asked) == sorted(mixed) and pair == (b1, b1)
    assert cand[(b2, b1)] == (2.86, 7.82) and cand[(b2, b2)] == (2.86, 7.70)
    mixed[(b2, b1)] = (2.60, 7.65)
    assert A.select(tf, tb, lambda f, b: mixed[(f, b)])[0] == (b2, b1)
    # with chunk counts: candidates within TIE of the fastest are equal and the
    # fewest chunks wins; one outside TIE never does
    nch = {(c256, c256): 28000, (c512, c512): 14458, (b1, b1): 2100, (b2, b2): 2900, (b2, b1): 2398,
           (b2, c512): 15000, (c512, b1): 1900}
    mixed[(b2, b1)] = (2.86, 7.65)
    pair, cand = A.select(tf, tb, lambda f, b: mixed[(f, b)], chunks=nch.get)
    assert sum(cand[(c512, b1)]) <= sum(cand[(b2, b1)]) * (1 + A.TIE) and pair == (c512, b1)
    nch[(c512, b1)] = 5000
    assert A.select(tf, tb, lambda f, b: mixed[(f, b)], chunks=nch.get)[0] == (b1, b1)
    nch[(c512, b1)] = 10
    slow = {k: (v[0] + 1.0, v[1] + 1.0) for k, v in mixed.items()}
    assert A.select(tf, tb, lambda f, b: slow[(f, b)], chunks=nch.get)[0] == (b1, b1)


def test_autotune_tune_plan_builds_every_candidate_it_compares(monkeypatch):
    """tune_plan with stand-in timings: every pair that select compares has
    its work list built (chunk counts known), including bins with two
    different multipliers, and the chosen pair's work list is returned."""
    from fbpinns_amd import autotune as A

    class Plan:
        def __init__(self, f, b):
            self.pair = (f, b)
            self.wl = type("W", (), {"n_chunks": {"chunks": 5000, "bins": 1000}[f[0]] + 10 * f[1] + b[1]})()

    built = []

    def build(f, b):
        built.append((f, b))
        return Plan(f, b)
    # 512-pair chunks fastest for both kernels (3.86 ms); 1x bins for both
    # within TIE of it (3.92 ms) with the fewest chunks
    times = {("chunks", 256): (1.00, 3.00), ("chunks", 512): (0.98, 2.88), ("bins", 1): (1.02, 2.90),
             ("bins", 2): (0.99, 2.95)}
    monkeypatch.setattr(A, "_time_launches", lambda net, th, dom, x, plan, reps: (
        times[plan.pair[0]][0], times[plan.pair[1]][1]))
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
    A._cache.clear()
    net = type("N", (), {"as_list": lambda self: [3, 64, 2, 7, 7], "nh": 2})()
    (fs, bs), plan = A.tune_plan(build, net, None, None, None, 864, 7_000_000)
    assert plan.pair == (fs, bs)
    assert (fs, bs) == (("bins", 1), ("bins", 1))
    assert (("bins", 2), ("bins", 1)) in built and (("bins", 2), ("chunks", 512)) in built
    assert not any("waves" in (f[0], b[0]) for f, b in built)
    A._cache.clear()


def test_autotune_wave_bins_default_and_candidate(monkeypatch):
    """One wave per chunk: the untimed default for a one-hidden-layer work
    list below MIN_PAIRS with at least WAVE_MIN_CHUNKS_PER_CU chunks per CU
    (never for two hidden layers), and a timed candidate of one-hidden-layer
    networks only, which wins when it is fastest."""
    from fbpinns_amd import autotune as A
    w = ("waves", A.WAVE_CHUNK)
    cut = A.WAVE_MIN_CHUNKS_PER_CU * 256
    assert A.default_spec(141_658, 1, 6078, 256) == w
    assert A.default_spec(30_000, 1, cut, 256) == w
    assert A.default_spec(30_000, 1, cut - 1, 256) == A.DEFAULT_SMALL
    # long chunks (here 100 pairs on average) keep a workgroup per chunk
    assert A.default_spec(100 * 6000, 1, 6000, 256) == A.DEFAULT_LARGE
    assert A.default_spec(100 * 1500, 1, 1500, 256) == A.DEFAULT_SMALL
    assert A.default_spec(A.WAVE_MAX_CHUNK_PAIRS * 1500, 1, 1500, 256) == w
    assert A.default_spec(400, 1, 15, 256) == A.DEFAULT_SMALL
    assert A.default_spec(141_658, 2, 6078, 256) == A.DEFAULT_SMALL
    assert A.default_spec(A.MIN_PAIRS, 1, 60_000, 256) == A.DEFAULT_LARGE
    assert w in A.candidates(1) and w not in A.candidates(2)

    class Plan:
        def __init__(self, f, b):
            self.pair = (f, b)
            self.wl = type("W", (), {"n_chunks": 1000 + 10 * f[1] + b[1]})()

    built = []

    def build(f, b):
        built.append((f, b))
        return Plan(f, b)
    net1 = type("N", (), {"as_list": lambda self: [1, 32, 1, 1, 1], "nh": 1})()
    # below MIN_PAIRS: default_spec with the caller's chunk and CU counts
    assert A.tune_plan(build, net1, None, None, None, 6000, 141_658, n_chunks=6078, n_cu=256)[0] == (w, w)
    assert A.tune_plan(build, net1, None, None, None, 15, 400, n_chunks=15, n_cu=256)[0] == (A.DEFAULT_SMALL,) * 2
    times = {("chunks", 256): (1.00, 3.00), ("chunks", 512): (0.98, 2.88), ("bins", 1): (1.02, 2.90),
             ("bins", 2): (0.99, 2.95), w: (0.70, 2.00)}
    monkeypatch.setattr(A, "_time_launches", lambda net, th, dom, x, plan, reps: (
        times[plan.pair[0]][0], times[plan.pair[1]][1]))
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
    A._cache.clear()
    built.clear()
    (fs, bs), plan = A.tune_plan(build, net1, None, None, None, 6000, 1_000_000)
    assert (fs, bs) == (w, w) and plan.pair == (w, w) and (w, w) in built
    A._cache.clear()


# ------------------------------------------------------------------ epilogue

def test_get_jmaps_matches_oracle():
    reqs = [((0, ()), (0, (0,)), (0, (0, 0))), ((0, (0, 0)), (0, (1, 1)), (0, (2, 2))),
            ((0, ()), (0, (0,)), (0, (1,)), (0, (0, 0))), ((0, ()),), ((0, (1,)), (0, (0,)))]
    for r in reqs:
        assert epilogue.get_jmaps(r) == om.get_jmaps(r)


def test_schema_masks_and_pick():
    assert epilogue.schema_masks([((0, ()), (0, (0,)), (0, (1,)), (0, (0, 0)))], 2) == (3, 1)
    assert epilogue.schema_masks([((0, (0, 0)), (0, (1, 1)), (0, (2, 2)))], 3) == (7, 7)
    with pytest.raises(FbpxError):
        epilogue.schema_masks([((0, (0, 1)),)], 2)
    with pytest.raises(FbpxError):
        epilogue.schema_masks([((0, (0, 0, 0)),)], 1)
    assert pick_schema(3, 7, 1) == (7, 3)
    assert pick_schema(2, 1, 0) == (3, 0)
    assert pick_schema(1, 0, 0) == (0, 0)


def _problem_params(prob, kw, xd):
    ap = {"static": {"problem": prob.init_params(**kw)[0]}, "trainable": {}}
    tr = prob.init_params(**kw)[1]
    if tr:
        ap["trainable"]["problem"] = {k: v.double() for k, v in tr.items()}
    return ap


class _Cubic(P.HarmonicOscillator1DHardBC):
    @staticmethod
    def constraining_fn(all_params, x_batch, u):
        return 1 + torch.tanh(x_batch[:, 0:1] / 0.1) ** 2 * (u + 0.2 * u ** 3)


@pytest.mark.parametrize("prob,kw,xd,req", [
    (P.HarmonicOscillator1DHardBC, dict(d=2, w0=20), 1, [((0, ()), (0, (0,)), (0, (0, 0)))]),
    (P.BurgersEquation2D, dict(), 2, [((0, ()), (0, (0,)), (0, (1,)), (0, (0, 0)))]),
    (P.WaveEquationConstantVelocity3D, dict(c0=1.3, source=P.wave_sources()), 3,
     [((0, (0, 0)), (0, (1, 1)), (0, (2, 2)))]),
    (P.PoissonMultiscale2D, dict(), 2, [((0, (0, 0)), (0, (1, 1)))]),
    (_Cubic, dict(d=2, w0=20), 1, [((0, ()), (0, (0,)), (0, (0, 0)))]),
])
def test_epilogue_affine_equals_generic(prob, kw, xd, req):
    """The Leibniz (affine) path and the nested-jvp (generic) path give the
    same u jets; the affine detection picks the right path."""
    torch.manual_seed(0)
    n = 64
    x = torch.rand((n, xd), dtype=torch.float64)
    d1, d2 = pick_schema(xd, *epilogue.schema_masks(req, xd))
    net = NetSpec.__new__(NetSpec)
    net.xd, net.d1mask, net.d2mask = xd, d1, d2
    net.J = 1 + bin(d1).count("1") + bin(d2).count("1")
    ap = _problem_params(prob, kw, xd)
    ep = epilogue.Epilogue(prob, ap, x, [(0, n)], [[x]], req, net)
    assert (ep.affine is not None) == (prob is not _Cubic)
    q = torch.randn((net.J, n), dtype=torch.float64)
    g = ep._ujs_generic(0, q, ap)
    if ep.affine is not None:
        # the precomputed jets of A and B with the Leibniz rule reproduce the
        # nested-jvp derivatives of constraining_fn (the rule the
        # fbpx_affine_jets kernel applies; its GPU output is checked against
        # the same generic path in tests/test_gpu_affine.py)
        A, B = ep.affine
        col = epilogue._col_index(net)
        for (iu, ixs), v in zip(req[0], g):
            if len(ixs) == 0:
                u = A[0] + B[0] * q[0]
            elif len(ixs) == 1:
                a_ = col[("d1", ixs[0])]
                u = A[a_] + B[a_] * q[0] + B[0] * q[a_]
            else:
                a_, b_ = col[("d1", ixs[0])], col[("d2", ixs[0])]
                u = A[b_] + B[b_] * q[0] + 2 * B[a_] * q[a_] + B[0] * q[b_]
            assert torch.allclose(u[:, None], v, rtol=1e-9, atol=1e-9)


# ------------------------------------------------------------------ library ABI

def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "fbpx.h")).read()
    declared = set(re.findall(r"\b(fbpx_[a-z_]+)\s*\(", hdr))
    assert len(declared) >= 15
    L = _lib.lib()
    for name in sorted(declared):
        assert hasattr(L, name), name
    assert set(_lib.SIGNATURES) == declared


def test_ctypes_signatures_match_header_parameter_counts():
    """Every ctypes argtypes list has as many entries as the header's
    declaration has parameters (a binding one argument short would pass
    garbage for the rest)."""
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "fbpx.h")).read(), flags=re.S)
    decls = re.findall(r"\b(?:int|int32_t|int64_t|const char\*)\s+(fbpx_\w+)\s*\(([^;]*?)\)\s*;", hdr, flags=re.S)
    assert len(decls) == len(_lib.SIGNATURES)
    for name, args in decls:
        n = 0 if args.strip() in ("", "void") else len(args.split(","))
        assert len(_lib.SIGNATURES[name][1]) == n, name


def test_library_shape_queries_and_error_paths():
    L = _lib.lib()
    assert L.fbpx_version().decode().startswith("fbpx")
    assert _lib.param_count(3, 64, 2) == 3 * 64 + 64 + 64 * 64 + 64 + 64 + 1
    assert _lib.param_count(1, 32, 1) == 97
    assert _lib.supported(3, 64, 2, 7, 7) and _lib.supported(1, 32, 1, 1, 1)
    assert not _lib.supported(3, 128, 2, 7, 7) and not _lib.supported(2, 32, 2, 1, 2)
    # argument validation returns an error code before touching the GPU
    rc = L.fbpx_inside_count(None, None, -1, 2, None, None, 3, None, None)
    assert rc == -1 and "bad sizes" in L.fbpx_last_error().decode()
    n = _lib.fbpx_net(2, 48, 2, 3, 3)
    rc = L.fbpx_forward(None, ctypes.byref(n), None, None, None, None, None)
    assert rc == -2 and "not compiled" in L.fbpx_last_error().decode()
    # one wave per bin: flags 0 / 1, one-hidden-layer networks only
    wl = _lib.fbpx_worklist()
    for nh, fw, bw, msg in ((2, 1, 0, "one-hidden-layer"), (2, 0, 1, "one-hidden-layer"),
                            (1, 2, 0, "must be 0 or 1"), (1, 0, -1, "must be 0 or 1")):
        wl.fwd_wave_bins, wl.bwd_wave_bins = fw, bw
        n = _lib.fbpx_net(1, 32, nh, 1, 1)
        rc = L.fbpx_forward(None, ctypes.byref(n), None, None, None, ctypes.byref(wl), None)
        assert rc == -1 and msg in L.fbpx_last_error().decode(), (nh, fw, bw)
    assert ctypes.sizeof(_lib.fbpx_worklist) == 7 * 4 + 4 + 9 * 8 + 2 * 4
    # the folded optimizer step: one hidden layer, every optimizer pointer given
    wl = _lib.fbpx_worklist()
    fake = ctypes.c_void_p(0x1000)
    for nh, mu, msg in ((2, fake, "one-hidden-layer"), (1, None, "null mu")):
        n = _lib.fbpx_net(1, 32, nh, 1, 1)
        rc = L.fbpx_backward_adam(None, ctypes.byref(n), fake, fake, fake, ctypes.byref(wl), None, fake, None, mu,
                                  fake, fake, fake, 1e-3, 0.9, 0.999, 1e-8)
        assert rc == -1 and msg in L.fbpx_last_error().decode(), nh
    # Adam over rows: one 32-bit element index per thread, n_rows * P < 2^31 - 255
    cnt = ctypes.c_void_p(0x2000)
    for P, n_rows in ((97, (2 ** 31 - 255) // 97 + 1), (1 << 16, 1 << 15)):
        rc = L.fbpx_adam_rows(None, fake, fake, fake, P, fake, fake, n_rows, cnt, None, 1e-3, 0.9, 0.999, 1e-8)
        assert rc == -1 and "bad sizes" in L.fbpx_last_error().decode(), (P, n_rows)


def test_pou_loss_and_residual_grad_validate_arguments():
    """fbpx_pou_loss and fbpx_residual_loss_grad reject bad arguments with an
    error code and message before any launch (no GPU needed).  The device
    pointers are never dereferenced: every case fails validation first."""
    L = _lib.lib()
    net = _lib.fbpx_net(1, 32, 1, 1, 1)
    fake = ctypes.c_void_p(0x1000)
    wl = _lib.fbpx_worklist()
    wl.n_points, wl.npou, wl.n_segs = 100, 1, 100

    def pou_loss(res, n_res, scratch_len, A=None, B=None):
        arr = (_lib.fbpx_residual * max(len(res), 1))(*res)
        return L.fbpx_pou_loss(None, ctypes.byref(net), ctypes.byref(wl), fake, fake, A, B, n_res, arr, fake, fake,
                               fake, scratch_len, fake, 0, None)

    def residual(**kw):
        r = _lib.fbpx_residual()
        r.offset, r.n, r.n_cols, r.n_terms, r.weight = 0, 100, 1, 1, 1.0
        r.coef, r.r = fake, fake
        r.jet[0] = 0
        for k, v in kw.items():
            setattr(r, k, v)
        return r
    ok = residual()
    cases = [
        (([ok], 0, 1), "n_res=0"),
        (([ok], _lib.FBPX_POU_LOSS_RES + 1, 1), "n_res"),
        (([ok], 1, 0), "scratch of 0 floats"),
        (([residual(n=101)], 1, 1), "residual 0: points"),
        (([residual(n_cols=_lib.FBPX_POU_LOSS_COLS + 1)], 1, 1), "columns"),
        (([residual(n_terms=0)], 1, 1), "terms"),
        (([residual(jet=(ctypes.c_int32 * 8)(3))], 1, 1), "jet column 3 of 3"),
        (([residual(jet=(ctypes.c_int32 * 8)(-1))], 1, 1), "memory column"),
        (([residual(fa=(ctypes.c_int32 * 8)(1))], 1, 1), "factors (1, 0) outside 1 columns"),
    ]
    for (res, n_res, sl), msg in cases:
        assert pou_loss(res, n_res, sl) == -1, msg
        assert msg in L.fbpx_last_error().decode(), (msg, L.fbpx_last_error().decode())
    assert pou_loss([ok], 1, 1, A=fake) == -1 and "both" in L.fbpx_last_error().decode()
    # the gradient form needs its gradient table
    fa = (ctypes.c_int32 * 1)(0)
    cols = (ctypes.c_void_p * 1)(0x1000)
    assert L.fbpx_residual_loss_grad(None, 10, 1, cols, None, 1, fa, fa, fake, 1.0, fake, fake, fake, 0, None,
                                     None) == -1
    assert "gcol" in L.fbpx_last_error().decode()

    # fbpx_residual_losses: memory columns only, offset 0, a scratch float per workgroup
    def losses(res, n_res, scratch_len):
        arr = (_lib.fbpx_residual * max(len(res), 1))(*res)
        return L.fbpx_residual_losses(None, n_res, arr, fake, scratch_len, fake, 0, None)
    mem = dict(jet=(ctypes.c_int32 * 8)(-1), col=(ctypes.c_void_p * 8)(0x1000))
    cases = [
        (([residual(**mem)], 0, 8), "n_res=0"),
        (([residual(**mem)] * 5, _lib.FBPX_POU_LOSS_RES + 1, 8), "n_res=5"),
        (([residual()], 1, 8), "jet 0 (columns here are device memory"),
        (([residual(offset=3, **mem)], 1, 8), "offset 3 (must be 0)"),
        (([residual(**mem), residual(n=0, **mem)], 2, 8), "residual 1: offset 0 (must be 0), 0 points"),
        (([residual(r=None, **mem)], 1, 8), "null coef / r"),
        (([residual(**mem), residual(**mem)], 2, 1), "scratch of 1 floats for 2 workgroups"),
        (([residual(fa=(ctypes.c_int32 * 8)(1), **mem)], 1, 8), "factors (1, 0) outside 1 columns"),
    ]
    for (res, n_res, sl), msg in cases:
        assert losses(res, n_res, sl) == -1, msg
        assert msg in L.fbpx_last_error().decode(), (msg, L.fbpx_last_error().decode())


def test_affine_jets_validates_requests():
    L = _lib.lib()
    n = _lib.fbpx_net(2, 32, 2, 3, 1)   # first derivatives along x0, x1; second along x0 only
    codes = (ctypes.c_int32 * 1)(4)      # d2/dx1^2: not carried
    ptrs = (ctypes.c_void_p * 1)(None)
    rc = L.fbpx_affine_jets(None, ctypes.byref(n), 8, 8, None, None, None, 1, codes, ptrs)
    assert rc == -1 and "not carried" in L.fbpx_last_error().decode()
    codes = (ctypes.c_int32 * 1)(0)
    rc = L.fbpx_affine_jets(None, ctypes.byref(n), 8, 4, None, None, None, 1, codes, ptrs)
    assert rc == -1 and "bad sizes" in L.fbpx_last_error().decode()
    rc = L.fbpx_affine_jets_backward(None, ctypes.byref(n), 8, 8, None, 0, codes, ptrs, None)
    assert rc == -1 and "bad sizes" in L.fbpx_last_error().decode()
    # n == 0 is a valid empty call
    assert L.fbpx_affine_jets(None, ctypes.byref(n), 0, 0, None, None, None, 1, codes, ptrs) == 0


def test_engine_refuses_to_run_without_gpu():
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from fbpinns_amd.engine import FusedEngine
    from fbpinns_amd.networks import FCN
    dec_cls, ap = setups.build((3,), (0.0,), (1.0,), 2.0, [1, 16, 1])
    with pytest.raises(FbpxError):
        FusedEngine(P.Problem, dec_cls, FCN, ap)


# ------------------------------------------------------------------ API surface

def test_decomposition_params_match_oracle():
    xs = [np.linspace(0, 1, 6), np.linspace(-1, 1, 4)]
    ws = get_subdomain_ws(xs, 1.9)
    dec = RectangularDecompositionND.init_params(xs, ws, (0.1, 2.0))[0]
    ref = ot.level_params(xs, ws, (0.1, 2.0))
    for a, b in zip(dec["subdomain"]["params"], ref[:6]):
        assert np.array_equal(a.numpy(), b.astype(np.float32))
    assert dec["m"] == 24 and dec["xd"] == 2


def test_schedulers():
    dec = RectangularDecompositionND.init_params([np.linspace(-1, 1, 5), np.linspace(0, 1, 4)],
                                                 get_subdomain_ws([np.linspace(-1, 1, 5), np.linspace(0, 1, 4)],
                                                                  2.0), (0, 1))[0]
    ap = {"static": {"decomposition": dec}}
    acts = list(S.AllActiveSchedulerND(ap, 5))
    assert np.all(acts[0] == 1) and all(a is None for a in acts[1:])
    sl = list(S.SlabSchedulerRectangularND(ap, 8, axis=1, steps_per_slab=2))
    assert sl[1] is None and sl[3] is None
    a0, a1 = sl[0].reshape(5, 4), sl[2].reshape(5, 4)
    assert np.all(a0[:, 0] == 1) and np.all(a0[:, 1:] == 0)
    assert np.all(a1[:, 0] == 2) and np.all(a1[:, 1] == 1) and np.all(a1[:, 2:] == 0)


def test_constants_guard_and_domain_samplers():
    c = Constants(n_steps=3)
    assert c.n_steps == 3 and c["seed"] == 0
    with pytest.raises(KeyError):
        Constants(not_a_key=1)
    ap = {"static": {"domain": RectangularDomainND.init_params(np.array([0.0, -1.0]), np.array([1.0, 1.0]))[0]}}
    g = RectangularDomainND.sample_interior(ap, None, "grid", (3, 2))
    assert np.allclose(g.numpy(), [[0, -1], [0, 1], [0.5, -1], [0.5, 1], [1, -1], [1, 1]])
    u = RectangularDomainND.sample_interior(ap, rng.PRNGKey(1), "uniform", (100, 1))
    assert u.shape == (100, 2) and (u[:, 0] >= 0).all() and (u[:, 1] >= -1).all() and (u[:, 1] <= 1).all()
    with pytest.raises(ValueError):
        RectangularDomainND.sample_interior(ap, None, "sobol", (3, 2))
