"""GPU: the full trainer loop (scheduler, work-list rebuilds, hipGraph
re-capture, rate / loss summaries, test L1, checkpoints) on small runs.

L1 checks are against the problems' exact solutions:
  * README HO config (HarmonicOscillator1D, d=2, w0=80, 15 subdomains of
    width 0.15, [1,32,1], 200 points, 20,000 steps, lr 1e-3): the reference
    figure (images/test-loss.png via README.md:18) shows FBPINN test L1 of
    about 2e-3 at the end; the fused engine reaches 1.57e-3 (its own seeded
    initialisation), bound L1 < 2.5e-3.
  * the hard-BC oscillator configuration at its stated 1,000 steps stays
    unconverged (L1 2.65e-1, L1 / std 0.754): held to 5 %, the kernels being
    deterministic.
  * matched L1: the plain per-op form and the fused engine, started from the
    same parameters, reach the same test L1 after 300 steps (rel 1e-3).
"""

import os

import numpy as np
import pytest
import torch

from fbpinns_amd import analysis, configs, problems, schedulers
from fbpinns_amd.constants import Constants
from fbpinns_amd.trainers import FBPINNTrainer, PINNTrainer

pytestmark = pytest.mark.gpu


def test_readme_ho_reaches_published_l1(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    c = configs.ho(problem=problems.HarmonicOscillator1D, problem_init_kwargs=dict(d=2, w0=80),
                   network_init_kwargs=dict(layer_sizes=[1, 32, 1]), n_steps=20000, summary_freq=5000,
                   test_freq=5000, model_save_freq=20000, run="ho_readme")
    tr = FBPINNTrainer(c)
    ap = tr.train()
    l1 = [r[4] for r in tr.u_test_losses]
    print("test L1 history", l1, flush=True)
    assert l1[-1] < l1[0]
    assert l1[-1] < 2.5e-3
    # checkpoint round trip in the reference layout
    c2, model = analysis.load_model("ho_readme", rootdir="results/")
    i, all_params, opt_states, active, losses = model
    assert i == 20000
    W = all_params["trainable"]["network"]["subdomain"]["layers"][0][0]
    assert tuple(W.shape) == (15, 32, 1)
    assert np.allclose(W.numpy(), ap["trainable"]["network"]["subdomain"]["layers"][0][0].cpu().numpy())
    count, mu, nu = opt_states[0]
    assert opt_states[1] == () and tuple(count.shape) == () and int(count) == 20000
    assert tuple(mu["network"]["subdomain"]["layers"][0][0].shape) == (15, 32, 1)
    x = torch.linspace(0, 1, 50)[:, None]
    u = analysis.FBPINN_solution(c2, all_params, np.ones(15, dtype=int), x)
    u_ref = tr.engine.model(x.cuda())
    assert torch.allclose(u, u_ref)


def test_hard_bc_config_1000_steps_l1():
    c = configs.ho()
    tr = FBPINNTrainer(c, write_outputs=False)
    tr.train()
    l1, l1n = tr.u_test_losses[-1][4], tr.u_test_losses[-1][5]
    print(f"ho hard BC, {c.n_steps} steps: L1 {l1:.4e} L1/std {l1n:.4f}", flush=True)
    assert abs(l1 - 0.265) <= 0.05 * 0.265 and abs(l1n - 0.754) <= 0.05 * 0.754


def test_checkpoint_resume_equals_uninterrupted_run():
    """30 steps, checkpoint (reference layout, read back with the safe
    loader), 20 more steps in a new engine built from the checkpoint: the
    parameters, the learnt coefficient and the Adam state equal those of 50
    uninterrupted steps bit for bit."""
    import pickle
    from fbpinns_amd.analysis import _resolve, _to_torch, safe_load
    from fbpinns_amd.trainers import checkpoint_model
    c = configs.ho(problem=problems.HarmonicOscillator1DInverse, problem_init_kwargs=dict(d=2, w0=20),
                   decomposition_init_kwargs=configs._rect((0.0,), (1.0,), (5,), 2.99),
                   network_init_kwargs=dict(layer_sizes=[1, 32, 1]), ns=((60,),))
    eng, sched, _, _ = FBPINNTrainer(c, write_outputs=False).setup(exact=False)
    active = next(iter(sched))
    eng.set_active(active)
    for _ in range(30):
        eng.step()
    torch.cuda.synchronize()
    blob = pickle.dumps(checkpoint_model(30, eng.all_params, eng.opt_state(), eng.active, []), protocol=4)
    for _ in range(20):
        eng.step()
    torch.cuda.synchronize()
    i, ap, ((count, mu, nu), empty), act, _ = _to_torch(_resolve(safe_load(blob)))
    assert i == 30 and int(count) == 30 and empty == ()
    eng2, *_ = FBPINNTrainer(c, write_outputs=False).setup(exact=False)
    for (W, b), (lW, lb) in zip(eng2.all_params["trainable"]["network"]["subdomain"]["layers"],
                                ap["trainable"]["network"]["subdomain"]["layers"]):
        W.copy_(lW)
        b.copy_(lb)
    eng2.all_params["trainable"]["problem"]["mu"].copy_(ap["trainable"]["problem"]["mu"])
    eng2.load_opt_state(count, mu, nu)
    eng2.set_active(act)
    for _ in range(20):
        eng2.step()
    torch.cuda.synchronize()
    print(f"[resume] mu {eng.pflat[0].item():.8f} vs {eng2.pflat[0].item():.8f}, count {int(eng2.count.item())}",
          flush=True)
    for a, b in ((eng.theta, eng2.theta), (eng.mu, eng2.mu), (eng.nu, eng2.nu), (eng.pflat, eng2.pflat),
                 (eng.pmu, eng2.pmu), (eng.pnu, eng2.pnu), (eng.count, eng2.count)):
        assert torch.equal(a, b)


def test_hard_bc_config_trains_and_matches_plain_l1():
    from fbpinns_amd.plain import PlainFBPINN
    c = configs.ho(n_steps=300)
    tr = FBPINNTrainer(c, write_outputs=False)
    eng, sched, x_test, u_exact = tr.setup()
    eng.set_active(next(iter(sched)))
    pl = PlainFBPINN(c.problem, c.decomposition, c.network, eng)
    pl.set_active(np.ones(eng.m, dtype=int))
    xt = x_test.cuda()
    ue = u_exact.cuda()
    l10 = torch.mean(torch.abs(ue - eng.model(xt))).item()
    for _ in range(300):
        eng.step()
        pl.step()
    l1f = torch.mean(torch.abs(ue - eng.model(xt))).item()
    l1p = torch.mean(torch.abs(ue - pl.model(xt))).item()
    print(f"L1 start {l10:.4e} fused {l1f:.6e} plain {l1p:.6e}", flush=True)
    assert l1f < l10
    assert abs(l1f - l1p) <= 1e-3 * l1p


def test_slab_scheduler_rebuilds_and_freezes(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    c = configs.burgers(ns=((100, 50),), n_test=(50, 20), n_steps=240,
                        scheduler_kwargs=dict(axis=1, steps_per_slab=40), summary_freq=40, test_freq=120,
                        model_save_freq=10 ** 9, run="burgers_small")
    tr = FBPINNTrainer(c, write_outputs=False)
    eng, sched, x_test, u_exact = tr.setup()
    snapshots = []
    for i, act in enumerate(sched):
        if act is not None:
            eng.set_active(act)
            snapshots.append((eng.active.copy(), eng.theta.clone()))
        loss = eng.step()
    torch.cuda.synchronize()
    assert np.isfinite(loss.item())
    assert len(snapshots) == 6
    # a slab fixed at change k keeps the parameters it had when it became fixed
    for (a_prev, th_prev), (a_next, th_next) in zip(snapshots[:-1], snapshots[1:]):
        newly_fixed = (a_prev == 1) & (a_next == 2)
        assert newly_fixed.any()
        final = eng.theta[torch.as_tensor(newly_fixed)]
        assert torch.equal(final, th_next[torch.as_tensor(newly_fixed)])
    assert int(eng.count.item()) == 240


def test_pinn_trainer_trains():
    c = Constants(problem=problems.HarmonicOscillator1DHardBC, problem_init_kwargs=dict(d=2, w0=20, sd=0.1),
                  network_init_kwargs=dict(layer_sizes=[1, 64, 64, 1]), n_steps=2000, ns=((100,),),
                  n_test=(200,), summary_freq=1000, test_freq=1000, model_save_freq=10 ** 9, run="pinn")
    tr = PINNTrainer(c, write_outputs=False)
    tr.train()
    l1 = [r[4] for r in tr.u_test_losses]
    print("PINN L1", l1, flush=True)
    assert tr.engine.m == 1 and l1[-1] < l1[0]
