"""Pins oracle/takes.py: hand-computed cases, a pure-Python brute force of the
inside matrix, and the reference's own take-order invariant
(decompositions_base.py:100-128: takes == row-major nonzero of the inside
matrix)."""

import numpy as np
import pytest

from oracle import takes as ot


def test_level_params_hand_case():
    xmins, xmaxs, wmins, wmaxs, flags, unnorms, pous = ot.level_params([np.array([0.0, 0.5, 1.0])],
                                                                       [np.array([0.6, 0.6, 0.6])], (0.5, 2.0))
    assert np.allclose(xmins[:, 0], [-0.3, 0.2, 0.7])
    assert np.allclose(xmaxs[:, 0], [0.3, 0.8, 1.3])
    # overlap widths: interior = xmax(i) - xmin(i+1); edges copy their neighbour side
    assert np.allclose(wmaxs[:, 0], [0.1, 0.1, 0.1])
    assert np.allclose(wmins[:, 0], [0.1, 0.1, 0.1])
    assert np.all(flags == 1) and np.all(pous == 0)
    assert np.allclose(unnorms, [[0.5, 2.0]] * 3)
    # one subdomain: window switched off (flag 0)
    _, _, _, _, f1, _, _ = ot.level_params([np.array([0.5])], [np.array([1.0])], (0.0, 1.0))
    assert f1[0, 0] == 0
    with pytest.raises(ValueError):
        ot.level_params([np.array([0.0, 1.0])], [np.array([0.4, 0.4])], (0.0, 1.0))


def _brute(x, lo, hi):
    n_take, m_take = [], []
    for i in range(x.shape[0]):
        for j in range(lo.shape[0]):
            if all(lo[j, d] <= x[i, d] <= hi[j, d] for d in range(x.shape[1])):
                n_take.append(i)
                m_take.append(j)
    return np.array(n_take, dtype=np.int64), np.array(m_take, dtype=np.int64)


@pytest.mark.parametrize("xd,n,m,seed", [(1, 50, 7, 0), (2, 120, 30, 1), (3, 80, 25, 2)])
def test_inside_points_brute_force(xd, n, m, seed):
    r = np.random.default_rng(seed)
    c = r.uniform(0, 1, (m, xd)).astype(np.float32)
    w = r.uniform(0.05, 0.4, (m, xd)).astype(np.float32)
    lo, hi = (c - w).astype(np.float32), (c + w).astype(np.float32)
    x = r.uniform(-0.1, 1.1, (n, xd)).astype(np.float32)
    k = min(5, m)
    x[:k] = lo[:k]           # points exactly on box edges are inside (inclusive)
    x[k:2 * k] = hi[:k]
    n_take, m_take, ims = ot.inside_points(x, lo, hi)
    bn, bm = _brute(x, lo, hi)
    assert np.array_equal(n_take, bn) and np.array_equal(m_take, bm)
    assert np.array_equal(ims, np.unique(bm))
    ips, d = ot.inside_models(x, lo, hi, np.arange(0, m, 2))
    bsel = np.isin(bm, np.arange(0, m, 2))
    assert np.array_equal(ips, np.unique(bn[bsel]))
    counts = np.array([(bm == j).sum() for j in range(0, m, 2)], dtype=float)
    assert np.isclose(d, counts.mean() ** (1 / xd))


def test_reference_selftest_invariant():
    """decompositions_base.py:108-128 setup: 10,000 points in [0,2]^2, 1,000
    boxes of half-width 0.1 centred in [1,3]^2; takes must equal np.nonzero."""
    r = np.random.default_rng(0)
    x = r.uniform(0, 2, (10000, 2)).astype(np.float32)
    c = r.uniform(1, 3, (1000, 2)).astype(np.float32)
    lo, hi = c - np.float32(0.1), c + np.float32(0.1)
    n_take, m_take, ims = ot.inside_points(x, lo, hi)
    ins = (x[:, None, :] >= lo[None]) & (x[:, None, :] <= hi[None])
    tn, tm = np.nonzero(ins.all(-1))
    assert np.array_equal(n_take, tn) and np.array_equal(m_take, tm)
    assert np.array_equal(ims, np.nonzero(ins.all(-1).any(0))[0])


def test_get_inputs_active_semantics():
    # boxes: 0=[0,1], 1=[0.5,1.5], 2=[1,2], 3=[5,6] (no points)
    lo = np.array([[0.0], [0.5], [1.0], [5.0]], np.float32)
    hi = np.array([[1.0], [1.5], [2.0], [6.0]], np.float32)
    x = np.array([[0.25], [0.75], [1.25], [1.75]], np.float32)
    pou = np.zeros(4)
    # model 1 inactive (0) but overlapping -> trained; model 2 fixed; model 3 fixed without points -> dropped
    takes, all_ims, active, active_ims, fixed_ims = ot.get_inputs(x, np.array([1, 0, 2, 2]), lo, hi, pou)
    assert list(active) == [1, 1, 2, 0]
    assert list(active_ims) == [0, 1] and list(fixed_ims) == [2]
    assert list(all_ims) == [0, 1, 2]
    m_take, n_take, p_take, np_take, npou = takes
    assert list(n_take) == [0, 1, 1, 2, 2, 3]
    assert list(all_ims[m_take]) == [0, 0, 1, 1, 2, 2]
    assert list(np_take) == [0, 1, 2, 3] and list(p_take) == [0, 1, 1, 2, 2, 3] and npou == 1


def test_get_inputs_multilevel_segments():
    lo = np.array([[0.0], [0.5], [-1.0]], np.float32)
    hi = np.array([[1.0], [1.5], [3.0]], np.float32)
    x = np.array([[0.25], [0.75], [1.25]], np.float32)
    pou = np.array([0, 0, 1])
    takes, all_ims, *_ = ot.get_inputs(x, np.ones(3, int), lo, hi, pou)
    m_take, n_take, p_take, np_take, npou = takes
    rows = np.stack([n_take, pou[all_ims][m_take]], 1)
    npu = np.unique(rows, axis=0)
    assert np.array_equal(npu[p_take], rows)          # unique round trip
    assert list(np_take) == [0, 0, 1, 1, 2, 2] and npou == 2


def test_split_takes_recovers_global_pairs():
    r = np.random.default_rng(4)
    lo = np.sort(r.uniform(0, 1, (8, 1)), 0).astype(np.float32)
    hi = (lo + 0.3).astype(np.float32)
    xg = r.uniform(0, 1.2, (60, 1)).astype(np.float32)
    offs, sizes = [0, 40], [40, 20]
    tips, local, cfs, cips = ot.get_x_batch(np.ones(8, int), xg, offs, sizes, lo, hi)
    x = xg[tips]
    takes, all_ims, *_ = ot.get_inputs(x, np.ones(8, int), lo, hi, np.zeros(8))
    takess = ot.split_takes(takes, cfs, cips, len(tips))
    m_take, n_take = takes[0], takes[1]
    for ic, (tk, ips) in enumerate(zip(takess, cips)):
        got = set(zip(ips[tk[1]].tolist(), tk[0].tolist()))
        sel = np.isin(n_take, ips)
        want = set(zip(n_take[sel].tolist(), m_take[sel].tolist()))
        assert got == want
        assert len(tk[3]) == len(np.unique(tk[1]))
