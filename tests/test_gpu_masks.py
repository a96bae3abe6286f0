"""GPU parity of the train step under scheduler masks (trainers.py:252-256,
347-357, 482-507, 531, 640), against the float64 oracle run under the same
mask:

  * inactive subdomains (0) that overlap the training points of the active
    ones are trained (active becomes 1, trainers.py:351-357);
  * fixed subdomains (2) with points add their u * w to the forward but get
    no gradient and no Adam update (trainers.py:252-256);
  * fixed subdomains without points are dropped (trainers.py:354-357);
  * inactive subdomains without points are unused.

Loss and the trained rows' gradients within 1e-4 relative of the oracle;
every untrained row's gradient exactly zero; the engine's remapped active
vector equal to the oracle's.  Adam: moments, parameters and the shared count
after steps across two set_active calls equal oracle/adam.py applied to the
kernels' gradients on the trained rows only.
"""

import numpy as np
import pytest
import torch

from fbpinns_amd.schedulers import SlabSchedulerRectangularND

from oracle import adam as oadam
from test_gpu_step import CASES, TOL, _build, _check_layer_grads, _oracle_loss_grad

pytestmark = pytest.mark.gpu


def _slab_states(ap, axis):
    s = SlabSchedulerRectangularND(ap, 10 ** 6, axis=axis, steps_per_slab=1)
    states = []
    for a in s:
        if a is None or len(states) == s.n_slabs:
            break
        states.append(a)
    return states


def _columns(ap, axis):
    c = np.array(ap["static"]["decomposition"]["xmins0"])[:, axis] + \
        np.array(ap["static"]["decomposition"]["xmaxs0"])[:, axis]
    lv = np.unique(np.round(c, 12))
    return np.searchsorted(lv, np.round(c, 12))


def _custom_burgers_state(ap):
    """x-columns of the 6 x 3 Burgers decomposition: column 5 active, columns
    3 and 0 fixed, the rest inactive.  Column 4 overlaps column 5's points
    (trained), column 3 holds some of them (stays fixed), column 0 holds none
    (dropped), columns 1-2 hold none (unused)."""
    col = _columns(ap, 0)
    a = np.zeros(len(col), dtype=int)
    a[col == 5] = 1
    a[col == 3] = 2
    a[col == 0] = 2
    return a


def _states(name, eng):
    ap = eng.all_params
    if name == "burgers_small":
        return _slab_states(ap, 1) + [_custom_burgers_state(ap)]
    return _slab_states(ap, 2)


@pytest.mark.parametrize("name", ["burgers_small", "wave_small"])
def test_masked_loss_and_gradients_match_oracle(name):
    eng, ap, constraints, ofac, has_mu = _build(name)
    states = _states(name, eng)
    assert len(states) >= 2
    seen = set()
    for k, state in enumerate(states):
        eng.set_active(state)
        loss, g, _ = eng.loss_and_grad()
        torch.cuda.synchronize()
        oloss, ogl, _, all_ims, oact, n_act = _oracle_loss_grad(eng, constraints, ofac, has_mu, active=state)
        kinds = dict(trained_from_inactive=int(((state == 0) & (oact == 1)).sum()),
                     fixed_kept=int(((state == 2) & (oact == 2)).sum()),
                     fixed_dropped=int(((state == 2) & (oact == 0)).sum()),
                     unused=int(((state == 0) & (oact == 0)).sum()))
        seen |= {key for key, v in kinds.items() if v}
        print(f"[mask] {name} state {k}: trained {n_act} of {len(all_ims)} used / {eng.m}; {kinds}; "
              f"pairs {eng.plan.wl.n_pairs}", flush=True)
        assert np.array_equal(np.asarray(eng.active), oact)
        el = abs(loss.item() - oloss.item()) / abs(oloss.item())
        print(f"  loss gpu {loss.item():.8e} oracle {oloss.item():.8e} rel {el:.2e}", flush=True)
        assert el <= TOL
        _check_layer_grads(g, [(gW[:n_act], gb[:n_act]) for gW, gb in ogl], all_ims[:n_act])
        untrained = np.setdiff1d(np.arange(eng.m), all_ims[:n_act])
        assert (g.detach().cpu()[torch.as_tensor(untrained)] == 0).all()
    if name == "burgers_small":
        assert seen == {"trained_from_inactive", "fixed_kept", "fixed_dropped", "unused"}, seen
    else:
        assert {"trained_from_inactive", "fixed_kept"} <= seen, seen


def _adam_oracle(th, mu, nu, g, rows, count, eng):
    """oracle/adam.py (float32) on the trained rows only."""
    p, m_, v_, c = oadam.adam_update(th[rows], mu[rows], nu[rows], g[rows], count, lr=eng.lr, b1=eng.b1,
                                     b2=eng.b2, eps=eng.eps)
    th, mu, nu = th.copy(), mu.copy(), nu.copy()
    th[rows], mu[rows], nu[rows] = p, m_, v_
    return th, mu, nu, c


@pytest.mark.parametrize("name", ["burgers_small", "wave_small"])
def test_adam_across_set_active_matches_oracle(name):
    """Two steps under the first slab state, two under the next: graph-replayed
    steps advance theta, mu, nu of the trained rows exactly as oracle Adam on
    the kernels' gradients, keep every other row (and a fixed row's moments
    from when it was trained), and count every step."""
    eng, *_ = _build(name, graph=True)
    states = _states(name, eng)
    plan_states = [states[0], states[-1]]
    th = eng.theta.detach().cpu().numpy().copy()
    mu = np.zeros_like(th)
    nu = np.zeros_like(th)
    count = 0
    for k, state in enumerate(plan_states):
        eng.set_active(state)
        rows = np.flatnonzero(np.asarray(eng.active) == 1)
        for s in range(2):
            _, g, _ = eng.loss_and_grad()
            eng.step()
            torch.cuda.synchronize()
            th, mu, nu, count = _adam_oracle(th, mu, nu, g.detach().cpu().numpy(), rows, count, eng)
            e_th = np.abs(eng.theta.detach().cpu().numpy() - th).max()
            e_mu = np.abs(eng.mu.detach().cpu().numpy() - mu).max() / max(np.abs(mu).max(), 1e-30)
            e_nu = np.abs(eng.nu.detach().cpu().numpy() - nu).max() / max(np.abs(nu).max(), 1e-30)
            print(f"[adam mask] {name} state {k} step {s}: {len(rows)} trained rows, count {count}, "
                  f"theta abs {e_th:.2e}, mu rel {e_mu:.2e}, nu rel {e_nu:.2e}", flush=True)
            assert int(eng.count.item()) == count
            assert e_th <= 1e-6 and e_mu <= 1e-5 and e_nu <= 1e-5
    # rows trained in the first state and fixed in the second kept their moments
    first = np.asarray(plan_states[0]) == 1
    fixed_later = (np.asarray(eng.active) == 2) & first
    if fixed_later.any():
        assert (np.abs(eng.mu.detach().cpu().numpy()[fixed_later]).max(1) > 0).all()
