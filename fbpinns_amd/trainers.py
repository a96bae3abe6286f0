"""Trainers (reference: trainers.py, trainers_base.py).

FBPINNTrainer(c).train() keeps the reference's flow (trainers.py:577-687):
parameter initialisation, constraint sampling, the scheduler loop with a new
work list on every scheduler change, rate / loss summaries, test L1 and
model checkpoints.  The train step is the fused engine (engine.FusedEngine);
there is no other implementation to fall back to.

PINNTrainer is the m = 1 special case: one subdomain whose box is the
domain, window flag 0 (window == 1) and the domain normalisation
(trainers.py:780-943 -> decompositions.py:169-170, domains.py:105-110).
"""

import os
import pickle
import time

import numpy as np
import torch

from fbpinns_amd.decompositions import RectangularDecompositionND
from fbpinns_amd.engine import FusedEngine
from fbpinns_amd.util import rng
from fbpinns_amd.util.logger import logger



def _stack_subdomain_params(plist):
    """Stack per-subdomain ({}, {"layers": [(W, b)]}) into leading-m arrays."""
    layers = [p[1]["layers"] for p in plist]
    return {"layers": [(torch.stack([l[i][0] for l in layers]), torch.stack([l[i][1] for l in layers]))
                       for i in range(len(layers[0]))]}


CALLABLE_TAG = "callable:"


def _to_plain(t):
    """numpy leaves in plain dicts / lists / tuples; a callable (e.g. a
    problem's c_fn) becomes the string "callable:<module>:<qualname>"."""
    if isinstance(t, dict):
        return {k: _to_plain(v) for k, v in t.items()}
    if isinstance(t, list):
        return [_to_plain(v) for v in t]
    if isinstance(t, tuple):
        return tuple(_to_plain(v) for v in t)
    if isinstance(t, torch.Tensor):
        return t.detach().cpu().numpy().copy()
    if callable(t):
        return f"{CALLABLE_TAG}{getattr(t, '__module__', '?')}:{getattr(t, '__qualname__', repr(t))}"
    return t


def checkpoint_model(i, all_params, opt_state, active, u_test_losses):
    """The reference's checkpoint tuple (trainers.py:721, pickled by
    trainers_base.py:64-69):
        (i, all_params, all_opt_states, active, u_test_losses)
    with numpy leaves in the reference's layout (W (m, out, in), b (m, out)).
    all_opt_states is optax.adam's state (ScaleByAdamState(count, mu, nu),
    EmptyState()) as plain tuples in the same positions: count an int32
    scalar, mu and nu mirroring all_params["trainable"] leaf for leaf.  Only
    dicts, lists, tuples, strings, numbers and numpy arrays, so that
    analysis.load_model reads it without running code from the file."""
    count, mu, nu = opt_state
    if isinstance(count, torch.Tensor):
        count = count.detach().cpu()
    return (int(i), _to_plain(all_params), ((np.asarray(count, dtype=np.int32).reshape(()), _to_plain(mu),
                                            _to_plain(nu)), ()),
            np.asarray(active, dtype=np.int64), np.asarray(u_test_losses, dtype=np.float64))


class _Trainer:
    """trainers_base.py:21-69 without tensorboard / figures (not part of the
    accelerated path): output directories, constants files, summaries and
    model pickles."""

    def __init__(self, c, write_outputs=True):
        self.write_outputs = write_outputs
        if write_outputs:
            c.get_outdirs()
            c.save_constants_file()
        logger.info(c)
        self.c = c

    def _print_summary(self, i, loss, rate, start):
        logger.info("[i: %i/%i] loss: %.4f rate: %.1f elapsed: %.2f hr %s" % (
            i, self.c.n_steps, loss, rate, (time.time() - start) / (60 * 60), self.c.run))

    def _save_model(self, i, model):
        if not self.write_outputs:
            return
        with open(self.c.model_out_dir + f"model_{i:08d}.jax", "wb") as f:
            pickle.dump(model, f, protocol=4)


class FBPINNTrainer(_Trainer):
    """FBPINN trainer (trainers.py:479-774)."""

    def _init_params(self, key):
        c = self.c
        all_params = {"static": {}, "trainable": {}}
        for tag, cl, kwargs in zip(["domain", "problem", "decomposition"], [c.domain, c.problem, c.decomposition],
                                   [c.domain_init_kwargs, c.problem_init_kwargs, c.decomposition_init_kwargs]):
            ps_ = cl.init_params(**kwargs)
            if ps_[0]:
                all_params["static"][tag] = ps_[0]
            if ps_[1]:
                all_params["trainable"][tag] = ps_[1]
        assert (all_params["static"]["domain"]["xd"] == all_params["static"]["problem"]["dims"][1]
                == all_params["static"]["decomposition"]["xd"])
        m = all_params["static"]["decomposition"]["m"]
        logger.info(f"Total number of subdomains: {m}")
        key, *subkeys = rng.split(key, m + 1)
        plist = [c.network.init_params(k, **c.network_init_kwargs) for k in subkeys]
        all_params["trainable"]["network"] = {"subdomain": _stack_subdomain_params(plist)}
        return key, all_params

    def _common_init(self, key, all_params, exact=True):
        """trainers.py:422-477: constraints, jet requests, test points, exact solution."""
        c = self.c
        for k in all_params["trainable"]:
            n = sum(v.numel() for v in _leaves(all_params["trainable"][k]))
            logger.info(f"\t{k}: {n:,}")
        key, subkey = rng.split(key)
        cg = c.problem.sample_constraints(all_params=all_params, domain=c.domain, key=subkey, sampler=c.sampler,
                                          batch_shapes=c.ns)
        for constraint_ in cg:
            for c_ in constraint_[:-1]:
                assert c_.shape[0] == constraint_[0].shape[0]
        required_ujss = [cst[-1] for cst in cg]
        constraints = [list(cst[:-1]) for cst in cg]
        logger.info(f"Total number of constraints: {len(constraints)}")
        x_test = c.domain.sample_interior(all_params=all_params, key=None, sampler="grid", batch_shape=c.n_test)
        u_exact = None
        if exact:
            logger.info("Computing exact solution..")
            try:
                u_exact = c.problem.exact_solution(all_params=all_params, x_batch=x_test, batch_shape=c.n_test)
            except NotImplementedError:
                u_exact = None
            logger.info("Computing done")
        return key, constraints, required_ujss, x_test, u_exact

    def _make_engine(self, all_params):
        c = self.c
        return FusedEngine(c.problem, c.decomposition, c.network, all_params, c.optimiser_kwargs,
                           device=c.device, graph=c.graph,
                           autotune=c.autotune, fused_loss=c.fused_loss, sort_points=c.sort_points)

    def setup(self, exact=True):
        """Parameters, constraints, test points and the engine (no steps)."""
        c = self.c
        key = rng.PRNGKey(c.seed)
        np.random.seed(c.seed)
        key, all_params = self._init_params(key)
        scheduler = c.scheduler(all_params=all_params, n_steps=c.n_steps, **c.scheduler_kwargs)
        key, constraints, required_ujss, x_test, u_exact = self._common_init(key, all_params, exact)
        eng = self._make_engine(all_params)
        self.engine = eng
        eng.set_constraints(constraints, required_ujss)
        return eng, scheduler, x_test, u_exact

    def train(self):
        c = self.c
        eng, scheduler, x_test, u_exact = self.setup()
        x_test_d = x_test.to(eng.device).contiguous()
        u_exact_d = u_exact.to(eng.device) if u_exact is not None else None

        self.u_test_losses = []
        start0 = start1 = time.time()
        report_time = 0.0
        lossval = None
        i = -1
        for i, active_ in enumerate(scheduler):
            if active_ is not None:
                startc = time.time()
                logger.info(f"[i: {i}/{c.n_steps}] Updating active inputs..")
                eng.set_active(active_)
                logger.info(f"[i: {i}/{c.n_steps}] Updating active inputs done ({time.time() - startc:.2f} s)")
                report_time += time.time() - startc
            if i == 0:
                start1, report_time = self._report(0, eng, lossval, start0, start1, report_time, x_test_d,
                                                   u_exact_d)
            lossval = eng.step()
            start1, report_time = self._report(i + 1, eng, lossval, start0, start1, report_time, x_test_d,
                                               u_exact_d)
        logger.info(f"[i: {i + 1}/{c.n_steps}] Training complete")
        torch.cuda.synchronize()
        return eng.all_params

    def _report(self, i, eng, lossval, start0, start1, report_time, x_test, u_exact):
        c = self.c
        summary_, test_, save_ = [(i % f == 0) for f in [c.summary_freq, c.test_freq, c.model_save_freq]]
        if summary_ or test_ or save_:
            if i != 0 and summary_:
                loss = lossval.item()
                rate = c.summary_freq / (time.time() - start1 - report_time)
                self._print_summary(i, loss, rate, start0)
                start1, report_time = time.time(), 0.0
            if test_ or save_:
                start2 = time.time()
                if test_ and u_exact is not None:
                    u_test = eng.model(x_test)
                    l1 = torch.mean(torch.abs(u_exact - u_test)).item()
                    l1n = l1 / u_exact.std().item()
                    self.u_test_losses.append([i, 0, 0, time.time() - start0, l1, l1n])
                    logger.info(f"[i: {i}/{c.n_steps}] test L1: {l1:.4e} (normalised {l1n:.4e})")
                if save_:
                    self._save_model(i, checkpoint_model(i, eng.all_params, eng.opt_state(), eng.active,
                                                         self.u_test_losses))
                report_time += time.time() - start2
        return start1, report_time


class PINNTrainer(FBPINNTrainer):
    """PINN trainer as a one-subdomain FBPINN (trainers.py:780-943)."""

    def _init_params(self, key):
        c = self.c
        all_params = {"static": {}, "trainable": {}}
        for tag, cl, kwargs in zip(["domain", "problem"], [c.domain, c.problem],
                                   [c.domain_init_kwargs, c.problem_init_kwargs]):
            ps_ = cl.init_params(**kwargs)
            if ps_[0]:
                all_params["static"][tag] = ps_[0]
            if ps_[1]:
                all_params["trainable"][tag] = ps_[1]
        d = all_params["static"]["domain"]
        xd = d["xd"]
        ps = RectangularDecompositionND._get_level_params(
            0, xd, [np.array([(a + b) / 2]) for a, b in zip(d["xmin"], d["xmax"])],
            [np.array([b - a]) for a, b in zip(d["xmin"], d["xmax"])], c.decomposition_init_kwargs["unnorm"])
        all_params["static"]["decomposition"] = RectangularDecompositionND._static(ps, xd)[0]
        key, subkey = rng.split(key)
        all_params["trainable"]["network"] = {
            "subdomain": _stack_subdomain_params([c.network.init_params(subkey, **c.network_init_kwargs)])}
        return key, all_params

    def _make_engine(self, all_params):
        c = self.c
        return FusedEngine(c.problem, RectangularDecompositionND, c.network, all_params, c.optimiser_kwargs,
                           device=c.device, graph=c.graph,
                           autotune=c.autotune, fused_loss=c.fused_loss, sort_points=c.sort_points)

    def train(self):
        from fbpinns_amd.schedulers import AllActiveSchedulerND
        sched = self.c.scheduler
        self.c.scheduler, kw = AllActiveSchedulerND, self.c.scheduler_kwargs
        self.c.scheduler_kwargs = {}
        try:
            return super().train()
        finally:
            self.c.scheduler, self.c.scheduler_kwargs = sched, kw


def _leaves(t):
    if isinstance(t, dict):
        for v in t.values():
            yield from _leaves(v)
    elif isinstance(t, (list, tuple)):
        for v in t:
            yield from _leaves(v)
    elif isinstance(t, torch.Tensor):
        yield t
