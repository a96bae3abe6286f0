"""Oracle: point -> subdomain take builder (numpy, integer arithmetic).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, by brute force over the full (n, m) inside matrix:
  * RectangularDecompositionND._get_level_params   decompositions.py:135-181
  * RectangularDecompositionND._inside_rectangleND decompositions.py:217-227
  * inside_points_batch / inside_models_batch      decompositions_base.py:70-82
    (the batched lax.map / lax.scan forms are equal to np.nonzero of the full
    inside matrix: that is the reference's own self-test,
    decompositions_base.py:100-128)
  * get_inputs                                    trainers.py:332-419
  * FBPINNTrainer._get_x_batch                    trainers.py:482-507
  * per-constraint take split                     trainers.py:544-571

Pinned by: the reference's own invariant (take order == row-major nonzero of
the inside matrix, decompositions_base.py:125-128) checked against an
independent per-point Python loop in tests/test_oracle_takes.py; hand-computed
tiny cases; np.unique(return_inverse) round-trips (npu[p_take] == rows).
"""

import numpy as np


def level_params(subdomain_xs, subdomain_ws, unnorm, il=0):
    """decompositions.py:135-181.  float64 numpy; returns f64 arrays
    [xmins, xmaxs, wmins, wmaxs, flags, unnorms, pous] each (m, .)."""
    xd = len(subdomain_xs)
    xs = np.stack(np.meshgrid(*subdomain_xs, indexing="ij"), 0)
    ws = np.stack(np.meshgrid(*subdomain_ws, indexing="ij"), 0)
    if xs.shape != ws.shape:
        raise ValueError("shape of subdomain_ws not same as subdomain_xs")
    xmins, xmaxs = xs - ws / 2, xs + ws / 2
    wmins, wmaxs = 0.5 * (xmaxs - xmins), 0.5 * (xmaxs - xmins)
    for i in range(xd):
        a = [slice(None)] * xd
        b = [slice(None)] * xd
        a[i] = slice(None, -1)
        b[i] = slice(1, None)
        a, b = (i,) + tuple(a), (i,) + tuple(b)
        wmaxs[a] = wmins[b] = xmaxs[a] - xmins[b]
        a = [slice(None)] * xd
        b = [slice(None)] * xd
        a[i] = 0
        b[i] = -1
        a, b = (i,) + tuple(a), (i,) + tuple(b)
        wmins[a] = wmaxs[a]
        wmaxs[b] = wmins[b]
    if (wmins <= 0).any() or (wmaxs <= 0).any():
        raise ValueError("some subdomains are not overlapping!")
    flat = lambda v: v.reshape(xd, -1).T
    xmins, xmaxs, wmins, wmaxs = flat(xmins), flat(xmaxs), flat(wmins), flat(wmaxs)
    s = (xmins.shape[0], 1)
    flags = np.zeros(s) if xmins.shape[0] == 1 else np.ones(s)
    unnorms = np.concatenate([unnorm[0] * np.ones(s), unnorm[1] * np.ones(s)], axis=1)
    pous = il * np.ones(s)
    return [xmins, xmaxs, wmins, wmaxs, flags, unnorms, pous]


def inside_matrix(x, xmins, xmaxs):
    """decompositions.py:217-227: inclusive box test in the arrays' dtype."""
    x = np.asarray(x)[:, None, :]
    return np.all((x >= np.asarray(xmins)[None]) & (x <= np.asarray(xmaxs)[None]), -1)


def inside_points(x, xmins, xmaxs):
    """decompositions.py:201-208 via decompositions_base.py:70-76.
    Returns n_take, m_take (row-major nonzero order) and inside_ims."""
    ins = inside_matrix(x, xmins, xmaxs)
    n_take, m_take = np.nonzero(ins)
    inside_ims = np.nonzero(ins.any(0))[0]
    return n_take.astype(np.int64), m_take.astype(np.int64), inside_ims


def inside_models(x, xmins, xmaxs, ims):
    """decompositions.py:210-215 via decompositions_base.py:78-82.
    Returns inside_ips (points inside >=1 model of ims) and d."""
    ims = np.asarray(ims, dtype=np.int64)
    ins = inside_matrix(x, np.asarray(xmins)[ims], np.asarray(xmaxs)[ims])
    inside_ips = np.nonzero(ins.any(1))[0]
    counts = ins.sum(0).astype(np.float64)
    d = counts.mean() ** (1 / np.asarray(x).shape[1]) if len(ims) else float("nan")
    return inside_ips, d


def get_inputs(x_batch, active, xmins, xmaxs, pou):
    """trainers.py:332-419.  `pou` is the (m,) integer pou index.
    Returns (m_take, n_take, p_take, np_take, npou), all_ims, active,
    active_ims, fixed_ims.  m_take indexes into all_ims = [active; fixed]."""
    m = len(pou)
    n_take, m_take, training_ims = inside_points(x_batch, xmins, xmaxs)
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
    pous = np.asarray(pou).astype(np.int64)[all_ims]
    npv = np.stack([n_take, pous[m_take]], axis=-1)
    npu, p_take = np.unique(npv, axis=0, return_inverse=True)
    p_take = p_take.reshape(-1)
    np_take = npu[:, 0]
    npou = len(np.unique(np.asarray(pou).astype(np.int64)))
    return (m_take, n_take, p_take, np_take, npou), all_ims, active, active_ims, fixed_ims


def get_x_batch(active, x_batch_global, constraint_offsets, constraint_sizes, xmins, xmaxs):
    """trainers.py:482-507 restricted to the index bookkeeping.
    Returns training_ips (into x_batch_global) and, per constraint, the
    indices of its surviving points relative to that constraint's own array,
    plus constraint_fs (n_train, c) and constraint_ips."""
    active = np.asarray(active)
    ims = np.arange(len(active))[active == 1]
    training_ips, _ = inside_models(x_batch_global, xmins, xmaxs, ims)
    n_glob = np.asarray(x_batch_global).shape[0]
    fs_global = np.zeros((n_glob, len(constraint_sizes)), dtype=bool)
    for ic, (o, s) in enumerate(zip(constraint_offsets, constraint_sizes)):
        fs_global[o:o + s, ic] = True
    constraint_fs = fs_global[training_ips]
    ix_ = np.arange(len(training_ips))
    constraint_ips = [ix_[f] for f in constraint_fs.T]
    local = [training_ips[constraint_ips[ic]] - constraint_offsets[ic] for ic in range(len(constraint_sizes))]
    return training_ips, local, constraint_fs, constraint_ips


def split_takes(takes, constraint_fs, constraint_ips, n_train):
    """trainers.py:544-571: per-constraint re-indexed takes."""
    m_take, n_take, p_take, np_take, npou = takes
    takess = []
    iu_ = np.arange(np_take.shape[0])
    for f, ips in zip(constraint_fs.T, constraint_ips):
        f1 = f
        f2 = f[np_take]
        ius = iu_[f2]
        f3 = f1[n_take]
        f4 = f2[p_take]
        inv = np.zeros(n_train, dtype=np.int64)
        inv[ips] = np.arange(len(ips))
        inv2 = np.zeros(np_take.shape[0], dtype=np.int64)
        inv2[ius] = np.arange(len(ius))
        takess.append((m_take[f3], inv[n_take[f3]], inv2[p_take[f4]], inv[np_take[f2]], npou))
    return takess
