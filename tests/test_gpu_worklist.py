"""GPU parity of the device take and work-list builder (fbpx_takes_count +
fbpx_build_worklist, fbpinns_amd.takes.device_plan):

  * takes (m_take, n_take, p_take, np_take, npou), all_ims and the remapped
    active vector bit-exact with the oracle's get_inputs (trainers.py:332-419)
    on the jet cases' geometries (1-D to 3-D, multilevel, one subdomain), under
    all-active and random {0, 1, 2} scheduler states;
  * every work-list array equal to the numpy restatement of the layout
    (tests/worklist_ref.py) for chunk caps and launch bins;
  * the reference's own take self-test geometry (decompositions_base.py:87-128:
    10,000 points in [0, 2)^2, 1,000 boxes of half-width 0.1 centred in
    [1, 3)^2, points outside every box) through fbpx_inside_* and the builder;
  * empty and degenerate inputs (no points, no pairs, one point);
  * the six slab states of the Burgers configuration, device builder against
    the host path, with both timed.
"""

import time

import numpy as np
import pytest
import torch

from fbpinns_amd import configs, ops
from fbpinns_amd.takes import device_plan, get_inputs
from fbpinns_amd.trainers import FBPINNTrainer

import setups
import worklist_ref
from oracle import takes as ot
from test_gpu_jets import CASES

pytestmark = pytest.mark.gpu

WL_KEYS = ("pair_point", "pair_seg", "chunk_model", "chunk_start", "point_seg_start", "seg_pair_start", "pair_slot",
           "fbin_chunk_start", "bbin_chunk_start")


def _boxes(ap):
    p = ap["static"]["decomposition"]["subdomain"]["params"]
    pou = ap["static"]["decomposition"]["subdomain"]["pou"].reshape(-1)
    return p[0], p[1], pou


def _ap_from_boxes(xmins, xmaxs, pou=None):
    m = xmins.shape[0]
    pou = np.zeros(m) if pou is None else pou
    return {"static": {"decomposition": {"m": m, "subdomain": {
        "params": [torch.tensor(xmins), torch.tensor(xmaxs)],
        "pou": torch.tensor(np.asarray(pou, dtype=np.float32).reshape(-1, 1))}}}}


def _check(ap, x, active, specs):
    """Device builder vs oracle takes and vs the numpy layout for each
    (chunk, fbins, bbins) in specs."""
    xmins, xmaxs, pou = _boxes(ap)
    lo, hi = xmins.cpu().numpy(), xmaxs.cpu().numpy()
    pou_np = pou.cpu().numpy()
    want_takes, want_all, want_act, want_aims, _ = ot.get_inputs(x.numpy(), active, lo, hi, pou_np)
    xd = x.cuda()
    for chunk, fb, bb in specs:
        plan, takes, all_ims, act = device_plan(xd, active, ap, chunk=chunk, fbins=fb, bbins=bb)
        got = takes.numpy()
        for a, b, nm in zip(got[:4], want_takes[:4], ["m_take", "n_take", "p_take", "np_take"]):
            assert np.array_equal(a, b), nm
        assert got[4] == want_takes[4]
        assert np.array_equal(all_ims, want_all) and np.array_equal(act, want_act)
        assert plan.n_active == len(want_aims)
        ref = worklist_ref.build_worklist(want_takes, want_all, len(want_aims), x.shape[0], "cpu", chunk=chunk,
                                          fbins=fb, bbins=bb)
        for k in WL_KEYS:
            assert np.array_equal(plan.wl.t[k].cpu().numpy(), ref.wl.t[k].numpy()), (k, chunk, fb, bb)
        assert plan.wl.size_list() == ref.wl.size_list(), (chunk, fb, bb)
        assert np.array_equal(plan.chunk_train.cpu().numpy(), ref.chunk_train.numpy())
        assert np.array_equal(plan.rows.cpu().numpy(), ref.rows.numpy())
        assert np.array_equal(plan.row_chunk_start.cpu().numpy(), ref.row_chunk_start.numpy())
    return len(want_takes[0])


SPECS = [(None, None, None), (512, None, None), (64, None, None), (None, 7, 5), (256, 13, 3), (None, 1, 1)]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_device_builder_matches_oracle_and_layout(case):
    name, nsub, lo, hi, ov, ls, _, (kind, npts), ml = case
    _, ap = setups.build(nsub, lo, hi, ov, ls, multilevel=ml)
    x = setups.grid_points(npts, lo, hi) if kind == "grid" else setups.uniform_points(npts, lo, hi, seed=7)
    m = ap["static"]["decomposition"]["m"]
    r = np.random.default_rng(3)
    states = [np.ones(m, dtype=int), r.integers(0, 3, m), r.integers(0, 3, m)]
    for k, st in enumerate(states):
        s = _check(ap, x, st, SPECS)
        print(f"[builder] {name} state {k}: {x.shape[0]} points, {s} pairs", flush=True)


def test_reference_self_test_geometry():
    """decompositions_base.py:87-128 with seeded numpy draws in place of jax.random."""
    r = np.random.default_rng(0)
    n, m = 10000, 1000
    x = r.uniform(0, 2, (n, 2)).astype(np.float32)
    c = r.uniform(1, 3, (m, 2)).astype(np.float32)
    xmins, xmaxs = c - np.float32(0.1), c + np.float32(0.1)
    # inside takes and the points / models inside (fbpx_inside_count / fill)
    nt, mt, counts = ops.inside_takes(torch.tensor(x).cuda(), torch.tensor(xmins).cuda(), torch.tensor(xmaxs).cuda())
    want = ot.inside_matrix(x, xmins, xmaxs)
    wn, wm = np.nonzero(want)
    assert np.array_equal(nt.cpu().numpy(), wn) and np.array_equal(mt.cpu().numpy(), wm)
    assert np.array_equal(np.nonzero(counts.cpu().numpy() > 0)[0], np.nonzero(want.any(1))[0])
    assert np.array_equal(np.unique(mt.cpu().numpy()), np.nonzero(want.any(0))[0])
    print(f"[self-test] {len(wn)} pairs, {int(want.any(1).sum())} of {n} points inside, "
          f"{int(want.any(0).sum())} of {m} boxes used", flush=True)
    # the whole builder on the same geometry, all active and a random state
    ap = _ap_from_boxes(xmins, xmaxs)
    for st in (np.ones(m, dtype=int), np.random.default_rng(1).integers(0, 3, m)):
        _check(ap, torch.tensor(x), st, [(None, None, None), (16, 9, 4)])


def test_degenerate_inputs():
    lo, hi = np.array([[0.0], [0.5]], np.float32), np.array([[0.6], [1.0]], np.float32)
    ap = _ap_from_boxes(lo, hi)
    cases = {"no points": np.zeros((0, 1), np.float32),
             "no pairs": np.array([[2.0], [-1.0], [np.nan]], np.float32),
             "one point": np.array([[0.55]], np.float32),
             "edges": np.array([[0.0], [0.5], [0.6], [1.0], [0.3]], np.float32)}
    for nm, x in cases.items():
        for st in (np.ones(2, dtype=int), np.array([2, 0]), np.array([0, 2])):
            s = _check(ap, torch.tensor(x), st, [(None, None, None), (1, 3, 2)])
            print(f"[degenerate] {nm} state {st.tolist()}: {s} pairs", flush=True)


def test_multilevel_pou_segments():
    """Three levels over the same points: segments are (point, level) rows."""
    from fbpinns_amd.decompositions import MultilevelRectangularDecompositionND
    xs = [[np.linspace(0, 1, 5)], [np.linspace(0, 1, 2)], [np.array([0.5])]]
    ws = [[np.full(5, 0.5)], [np.full(2, 1.2)], [np.array([1.3])]]
    dec = MultilevelRectangularDecompositionND.init_params(xs, ws, (0.0, 1.0))[0]
    ap = {"static": {"decomposition": dec}}
    x = setups.uniform_points(3000, (0.0,), (1.0,), seed=5)
    m = dec["m"]
    for st in (np.ones(m, dtype=int), np.random.default_rng(2).integers(0, 3, m)):
        _check(ap, x, st, [(None, None, None), (32, 5, 5)])


def test_burgers_slab_changes_timed():
    """The Burgers configuration's six slab states: device builder output
    equal to the host path's (get_inputs + the numpy layout), both timed."""
    c = configs.CONFIGS["burgers"]()
    tr = FBPINNTrainer(c, write_outputs=False)
    eng, sched, _, _ = tr.setup(exact=False)
    from fbpinns_amd.takes import get_x_batch
    states, seen = [], set()
    for a in sched:
        if a is None:
            continue
        key = tuple(np.asarray(a).tolist())
        if key in seen:
            continue
        seen.add(key)
        states.append(np.asarray(a))
        if len(states) == 6:
            break
    assert len(states) == 6
    t_dev = t_host = 0.0
    for k, st in enumerate(states):
        tips, _, _, _ = get_x_batch(st, eng.all_params, eng.x_global, eng.c_offsets, eng.c_sizes, eng.decomposition)
        x = eng.x_global[torch.as_tensor(tips, device=eng.device)].contiguous()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        plan, takes, all_ims, act = device_plan(x, st, eng.all_params, chunk=512)
        torch.cuda.synchronize()
        td = time.perf_counter() - t0
        t0 = time.perf_counter()
        htakes, hall, (hact, *_) = get_inputs(x, st, eng.all_params, eng.decomposition)
        ref = worklist_ref.build_worklist(htakes, hall, int((hact == 1).sum()), x.shape[0], "cpu", chunk=512)
        th = time.perf_counter() - t0
        for a, b in zip(takes.numpy()[:4], htakes[:4]):
            assert np.array_equal(a, b)
        for key in WL_KEYS:
            assert np.array_equal(plan.wl.t[key].cpu().numpy(), ref.wl.t[key].numpy()), key
        if k:   # the first call carries one-off initialisation
            t_dev += td
            t_host += th
        print(f"[slab {k}] {x.shape[0]} points, {plan.wl.n_pairs} pairs, {plan.n_active} trained / "
              f"{len(all_ims)} used: device {td * 1e3:.2f} ms, host {th * 1e3:.2f} ms", flush=True)
    print(f"[slabs 1-5] device {t_dev * 1e3:.2f} ms, host {t_host * 1e3:.2f} ms in all", flush=True)
