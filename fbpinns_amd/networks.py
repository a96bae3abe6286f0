"""Subdomain networks (reference: networks.py).

Only the tanh FCN is accelerated: its forward/backward jets run in the HIP
kernels (fbpinns_amd/csrc/fbpx_net.hpp).  The per-point `network_fn` below is
the API-level definition (networks.py:60-68) used when a single point is
evaluated through the plugin interface.
"""

import math

import numpy as np
import torch

from fbpinns_amd.util import rng


class Network:
    """Base network class (networks.py:14-34).  Static methods only."""

    @staticmethod
    def init_params(*args):
        """Returns ({static}, {trainable}) parameter dicts."""
        raise NotImplementedError

    @staticmethod
    def network_fn(params, x):
        """Forward model for a SINGLE point x of shape (xd,)."""
        raise NotImplementedError


class FCN(Network):
    """Fully connected tanh network (networks.py:39-68).

    init="uniform" (default) draws W, b ~ U(-sqrt(1/fan_in), +sqrt(1/fan_in))
    as networks.py:50-58; init="glorot_normal" draws W ~ N(0, 2/(fan_in+fan_out))
    and b = 0."""

    @staticmethod
    def init_params(key, layer_sizes, init="uniform"):
        keys = rng.split(key, len(layer_sizes) - 1)
        params = [FCN._random_layer_params(k, m, n, init) for k, m, n in zip(keys, layer_sizes[:-1], layer_sizes[1:])]
        return {}, {"layers": params}

    @staticmethod
    def _random_layer_params(key, m, n, init="uniform"):
        w_key, b_key = rng.split(key)
        if init == "uniform":
            v = math.sqrt(1 / m)
            w = rng.uniform(w_key, (n, m), -v, v)
            b = rng.uniform(b_key, (n,), -v, v)
        elif init == "glorot_normal":
            w = rng.normal(w_key, (n, m)) * math.sqrt(2.0 / (m + n))
            b = np.zeros((n,))
        else:
            raise ValueError(f"unknown init {init!r}")
        return torch.tensor(w, dtype=torch.float32), torch.tensor(b, dtype=torch.float32)

    @staticmethod
    def network_fn(params, x):
        layers = params["trainable"]["network"]["subdomain"]["layers"]
        for w, b in layers[:-1]:
            x = torch.tanh(w @ x + b)
        w, b = layers[-1]
        return w @ x + b


def norm(mu, sd, x):
    return (x - mu) / sd


def unnorm(mu, sd, x):
    return x * sd + mu
