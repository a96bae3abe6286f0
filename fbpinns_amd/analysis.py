"""Loading saved runs and evaluating trained models (reference: analysis.py).

Checkpoints are pickles written by the trainers at
results/models/<run>/model_{i:08d}.jax holding
(i, all_params, all_opt_states, active, u_test_losses) with numpy leaves in
the reference's layout (W (m, out, in), b (m, out)) (trainers_base.py:64-69,
trainers.py:721); all_opt_states is optax.adam's (ScaleByAdamState(count,
mu, nu), EmptyState()) as plain tuples in the same positions
(trainers.checkpoint_model).  They are read with an unpickler that only
rebuilds containers, numbers, strings and numpy arrays (nothing in the file
is run); callables saved as "callable:<module>:<qualname>" strings are
looked up again in modules that are already imported.
"""

import io
import os
import pickle
import sys

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


_SAFE_GLOBALS = {
    ("numpy._core.multiarray", "_reconstruct"), ("numpy.core.multiarray", "_reconstruct"),
    ("numpy._core.multiarray", "scalar"), ("numpy.core.multiarray", "scalar"),
    ("numpy", "ndarray"), ("numpy", "dtype"),
}


class _ArrayUnpickler(pickle.Unpickler):
    """Unpickler for checkpoints: numpy arrays and plain containers only."""

    def find_class(self, module, name):
        if (module, name) in _SAFE_GLOBALS:
            return super().find_class(module, name)
        raise pickle.UnpicklingError(f"checkpoint refers to {module}.{name}: only numpy arrays and plain "
                                     f"containers are loaded")


def safe_load(path_or_bytes):
    """Load a checkpoint pickle without running code from it."""
    if isinstance(path_or_bytes, (bytes, bytearray)):
        return _ArrayUnpickler(io.BytesIO(path_or_bytes)).load()
    with open(path_or_bytes, "rb") as f:
        return _ArrayUnpickler(f).load()


def _resolve(t):
    """Callables saved as "callable:<module>:<qualname>" from modules that
    are already imported (no import runs); other strings stay as they are."""
    from fbpinns_amd.trainers import CALLABLE_TAG
    if isinstance(t, dict):
        return {k: _resolve(v) for k, v in t.items()}
    if isinstance(t, list):
        return [_resolve(v) for v in t]
    if isinstance(t, tuple):
        return tuple(_resolve(v) for v in t)
    if isinstance(t, str) and t.startswith(CALLABLE_TAG):
        mod, _, qual = t[len(CALLABLE_TAG):].partition(":")
        obj = sys.modules.get(mod)
        for part in qual.split("."):
            obj = getattr(obj, part, None) if obj is not None else None
        if obj is None:
            logger.warning(f"checkpoint callable {mod}:{qual} is not available; left as its name")
            return t
        return obj
    return t


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
    return c, _to_torch(_resolve(safe_load(path)))


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
