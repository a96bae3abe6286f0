"""Problem domains (reference: domains.py).  Host-side samplers; points are
returned as float32 torch tensors on the CPU (the trainer moves them)."""

import numpy as np
import torch

from fbpinns_amd import networks
from fbpinns_amd.util import rng


class Domain:
    """Base domain class (domains.py:18-55)."""

    @staticmethod
    def init_params(*args):
        raise NotImplementedError

    @staticmethod
    def sample_interior(all_params, key, sampler, batch_shape):
        raise NotImplementedError

    @staticmethod
    def sample_boundaries(all_params, key, sampler, batch_shapes):
        raise NotImplementedError

    @staticmethod
    def norm_fn(all_params, x):
        """PINN-only normalisation of a SINGLE point (xd,)."""
        raise NotImplementedError


class RectangularDomainND(Domain):
    """domains.py:60-142."""

    @staticmethod
    def init_params(xmin, xmax):
        xmin, xmax = np.asarray(xmin, dtype=np.float64), np.asarray(xmax, dtype=np.float64)
        assert xmin.shape == xmax.shape
        assert xmin.ndim == 1
        return {"xd": len(xmin), "xmin": xmin, "xmax": xmax}, {}

    @staticmethod
    def sample_interior(all_params, key, sampler, batch_shape):
        d = all_params["static"]["domain"]
        return RectangularDomainND._rectangle_samplerND(key, sampler, d["xmin"], d["xmax"], batch_shape)

    @staticmethod
    def norm_fn(all_params, x):
        d = all_params["static"]["domain"]
        xmin = torch.as_tensor(d["xmin"], dtype=x.dtype, device=x.device)
        xmax = torch.as_tensor(d["xmax"], dtype=x.dtype, device=x.device)
        return networks.norm((xmax + xmin) / 2, (xmax - xmin) / 2, x)

    @staticmethod
    def _rectangle_samplerND(key, sampler, xmin, xmax, batch_shape):
        """domains.py:112-142 for the samplers the configurations use.  grid:
        per-axis linspace (computed in float64, rounded once to float32) on an
        ij meshgrid; uniform: xmin + (xmax - xmin) * s, s from the seeded key.
        The reference's quasi-random samplers (sobol, halton) and boundary
        sampling are not part of this engine and raise."""
        xmin, xmax = np.asarray(xmin, dtype=np.float64), np.asarray(xmax, dtype=np.float64)
        assert xmin.shape == xmax.shape and xmin.ndim == 1
        xd = len(xmin)
        assert len(batch_shape) == xd
        if sampler not in ("grid", "uniform"):
            raise ValueError(f"sampler {sampler!r} is not supported (grid, uniform)")
        if sampler == "grid":
            xs = [np.linspace(a, b, n) for a, b, n in zip(xmin, xmax, batch_shape)]
            xx = np.stack(np.meshgrid(*xs, indexing="ij"), -1).reshape((-1, xd))
            return torch.tensor(xx, dtype=torch.float32)
        n = int(np.prod(batch_shape))
        s = rng.uniform(key, (n, xd))
        x = xmin[None] + (xmax - xmin)[None] * s
        return torch.tensor(x, dtype=torch.float32)


class InstancesDomain1D(Domain):
    """n_inst copies of [0, 1] placed at x = spacing * i (independent problem
    instances sharing one launch).  Samplers return the points of every
    instance, instance-major."""

    @staticmethod
    def init_params(n_inst, spacing=2.0):
        return {"xd": 1, "n_inst": int(n_inst), "spacing": float(spacing),
                "xmin": np.array([0.0]), "xmax": np.array([spacing * (n_inst - 1) + 1.0])}, {}

    @staticmethod
    def sample_interior(all_params, key, sampler, batch_shape):
        d = all_params["static"]["domain"]
        loc = RectangularDomainND._rectangle_samplerND(key, sampler, np.array([0.0]), np.array([1.0]),
                                                       batch_shape).double()
        off = d["spacing"] * torch.arange(d["n_inst"], dtype=torch.float64)
        return (off[:, None, None] + loc[None]).reshape(-1, 1).float()

    @staticmethod
    def instance_of(all_params, x_batch):
        d = all_params["static"]["domain"]
        return torch.floor(x_batch[:, 0].double() / d["spacing"] + 1e-9).long()
