"""Splittable random keys.

The reference draws with jax.random (threefry keys: PRNGKey/split,
trainers.py:583,603,437; networks.py:44-57; domains.py:137).  jax is not part
of this engine, so keys here are numpy SeedSequences with the same
split-and-consume discipline; draws are therefore reproducible for a seed but
are not the reference's threefry values.
"""

import numpy as np


def PRNGKey(seed):
    return np.random.SeedSequence(int(seed))


def split(key, num=2):
    return key.spawn(num)


def _gen(key):
    if key is None:
        raise ValueError("a random key is required")
    return np.random.default_rng(key)


def uniform(key, shape, minval=0.0, maxval=1.0):
    """float64 draws in [minval, maxval)."""
    return _gen(key).uniform(minval, maxval, size=shape)


def normal(key, shape):
    return _gen(key).standard_normal(size=shape)
