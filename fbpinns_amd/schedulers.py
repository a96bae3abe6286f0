"""Active-subdomain schedulers (reference: schedulers.py).

Iterables yielding, per training step, either None (unchanged) or an int
array `active` of length m with 0 = inactive (trained if it overlaps the
training points), 1 = active, 2 = fixed.  Host-side; the engine turns a new
`active` into a new work list (masked gradients and Adam updates).
"""

import numpy as np


class ActiveScheduler:
    """Base scheduler (schedulers.py:16-38)."""

    def __init__(self, all_params, n_steps):
        self.n_steps = n_steps
        self.m = all_params["static"]["decomposition"]["m"]
        self.xd = all_params["static"]["decomposition"]["xd"]

    def __len__(self):
        return self.n_steps

    def __iter__(self):
        raise NotImplementedError


class AllActiveSchedulerND(ActiveScheduler):
    """Every subdomain active for every step (schedulers.py:41-49)."""

    def __iter__(self):
        for i in range(self.n_steps):
            yield np.ones(self.m, dtype=int) if i == 0 else None


class SlabSchedulerRectangularND(ActiveScheduler):
    """Time-marching slabs: the subdomain grid is cut along `axis` into slabs
    of `per_slab` subdomain layers (default 1); slab k is active for
    `steps_per_slab` steps, slabs < k are fixed, slabs > k inactive.  Follows
    the reference schedulers' state semantics (schedulers.py:117-135: 0
    inactive, 1 active, 2 fixed) with a fixed step count per slab.  The
    reference's point / line / plane schedulers (schedulers.py:52-161) are not
    part of this engine; a user scheduler yielding the same states drives it
    the same way."""

    def __init__(self, all_params, n_steps, axis, steps_per_slab, per_slab=1):
        super().__init__(all_params, n_steps)
        c = np.array(all_params["static"]["decomposition"]["xmins0"])[:, axis] + \
            np.array(all_params["static"]["decomposition"]["xmaxs0"])[:, axis]
        levels = np.unique(np.round(c, 12))
        self.layer = np.searchsorted(levels, np.round(c, 12)) // per_slab
        self.n_slabs = int(self.layer.max()) + 1
        self.steps_per_slab = int(steps_per_slab)

    def __iter__(self):
        for i in range(self.n_steps):
            if i % self.steps_per_slab == 0:
                k = min(i // self.steps_per_slab, self.n_slabs - 1)
                if i == 0 or i // self.steps_per_slab < self.n_slabs:
                    active = np.where(self.layer < k, 2, np.where(self.layer == k, 1, 0)).astype(int)
                    yield active
                    continue
            yield None
