"""Loading saved runs and evaluating trained models (reference: analysis.py).

Checkpoints are pickles written by the trainers at
results/models/<run>/model_{i:08d}.jax holding
(i, all_params, all_opt_states, active, u_test_losses) with numpy leaves in
the reference's layout (W (m, out, in), b (m, out)) (trainers_base.py:64-69,
trainers.py:721).  The optimiser state is (ScaleByAdamState(count, mu, nu),
EmptyState()) as named tuples of this package.
"""

import os
import pickle

import numpy as np
import torch

from fbpinns_amd.util.logger import logger


class DictToObj:
    """Attribute view of a constants dict (util/other.py DictToObj)."""

    def __init__(self, copy=False, **kwargs):
        for k, v in kwargs.items():
            setattr(self, k, v)

    def __getitem__(self, k):
        return getattr(self, k)


def _to_torch(t):
    if isinstance(t, dict):
        return {k: _to_torch(v) for k, v in t.items()}
    if isinstance(t, list):
        return [_to_torch(v) for v in t]
    if isinstance(t, tuple) and not hasattr(t, "_fields"):
        return tuple(_to_torch(v) for v in t)
    if isinstance(t, np.ndarray) and t.dtype != object:
        return torch.from_numpy(t)
    return t


def load_model(run, i=None, rootdir="results/"):
    """analysis.py:18-46: constants object and model tuple of a saved run
    (the latest checkpoint unless i is given)."""
    model_dir = rootdir + f"models/{run}/"
    summary_dir = rootdir + f"summaries/{run}/"
    with open(summary_dir + f"constants_{run}.pickle", "rb") as f:
        c_dict = pickle.load(f)
    c = DictToObj(**c_dict)
    if i is None:
        last = sorted(os.listdir(model_dir))[-1]
        i = int(os.path.splitext(last)[0].split("_")[1])
    path = model_dir + f"model_{i:08d}.jax"
    logger.info(f"Loading model from:\n{path}")
    with open(path, "rb") as f:
        model = pickle.load(f)
    return c, _to_torch(model)


def FBPINN_solution(c, all_params, active, x_batch, device="cuda:0"):
    """analysis.py:49-58: the trained FBPINN's solution at x_batch, evaluated
    with the fused forward kernels (value jets only)."""
    from fbpinns_amd.engine import FusedEngine
    eng = FusedEngine(c.problem, c.decomposition, c.network, all_params, device=device, graph=False)
    x = torch.as_tensor(x_batch, dtype=torch.float32).to(eng.device)
    return eng.model(x, active)


def PINN_solution(c, all_params, x_batch, device="cuda:0"):
    """analysis.py:60-71 (a PINN is saved as a one-subdomain FBPINN)."""
    from fbpinns_amd.decompositions import RectangularDecompositionND
    from fbpinns_amd.engine import FusedEngine
    eng = FusedEngine(c.problem, RectangularDecompositionND, c.network, all_params, device=device, graph=False)
    x = torch.as_tensor(x_batch, dtype=torch.float32).to(eng.device)
    return eng.model(x)
