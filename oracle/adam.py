"""Oracle: optax.adam update (numpy).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

optax is a third-party dependency of the reference (pyproject.toml:21,
`optax>=0.1.4`, unpinned, not vendored); the reference calls it at
trainers.py:430 (optax.adam(**optimiser_kwargs)) and trainers.py:294-295
(update + apply_updates).  Published algorithm restated (optax
scale_by_adam + scale_by_learning_rate + apply_updates, nesterov=False,
eps_root=0):
    mu    = (1-b1) * g   + b1 * mu
    nu    = (1-b2) * g^2 + b2 * nu
    count = count + 1                      (int32, shared by all leaves)
    mhat  = mu / (1 - b1^count)
    vhat  = nu / (1 - b2^count)
    p     = p + (-lr) * (mhat / (sqrt(vhat + eps_root) + eps))

Pinned by tests/test_oracle_adam.py: Kingma & Ba Algorithm 1 closed forms
(first step moves every coordinate by -lr*sign(g) up to eps; a constant
gradient gives exactly -lr*g/(|g|+eps) per step), and an independent scalar
Python loop in float64.
"""

import numpy as np


def adam_update(p, mu, nu, g, count, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, eps_root=0.0):
    """One optax.adam step.  Arrays keep their dtype; returns (p, mu, nu, count+1)."""
    dt = p.dtype.type
    mu = dt(1 - b1) * g + dt(b1) * mu
    nu = dt(1 - b2) * (g ** 2) + dt(b2) * nu
    c = count + 1
    bc1 = dt(1) - dt(b1) ** dt(c)
    bc2 = dt(1) - dt(b2) ** dt(c)
    mhat = mu / bc1
    vhat = nu / bc2
    upd = mhat / (np.sqrt(vhat + dt(eps_root)) + dt(eps))
    p = p + dt(-lr) * upd
    return p, mu, nu, c
