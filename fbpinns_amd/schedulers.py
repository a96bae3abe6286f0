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


class _SubspacePointSchedulerRectangularND(ActiveScheduler):
    """Activates subdomains as a radius grows from a point in a subspace
    (schedulers.py:52-135)."""

    def __init__(self, all_params, n_steps, point, iaxes):
        super().__init__(all_params, n_steps)
        point = np.array(point)
        iaxes = list(iaxes)
        if point.ndim != 1:
            raise Exception("ERROR: point.ndim != 1")
        if len(point) > self.xd:
            raise Exception("ERROR: len(point) > self.xd")
        if len(iaxes) + len(point) != self.xd:
            raise Exception("ERROR: len(iaxes) + len(point) != self.xd")
        self.point, self.iaxes = point, iaxes
        self.xmins0 = np.array(all_params["static"]["decomposition"]["xmins0"]).copy()
        self.xmaxs0 = np.array(all_params["static"]["decomposition"]["xmaxs0"]).copy()

    @staticmethod
    def _get_radii(point, xmins, xmaxs):
        """Shortest and longest distance from point to each box (schedulers.py:74-103)."""
        assert xmins.shape[1] == xmaxs.shape[1] == point.shape[0]
        point = point[None]
        c_inside = np.prod((point >= xmins) & (point <= xmaxs), axis=1).astype(bool)
        pmin = np.clip(point, xmins, xmaxs)
        ds = np.stack([point - xmins, point - xmaxs], axis=0)
        i = np.argmax(np.abs(ds), axis=0)[None]
        pmax = point - np.take_along_axis(ds, i, axis=0)[0]
        rmin = np.sqrt(np.sum((pmin - point) ** 2, axis=1))
        rmax = np.sqrt(np.sum((pmax - point) ** 2, axis=1))
        rmin[c_inside] = 0.0
        return rmin, rmax

    def __iter__(self):
        ic = [i for i in range(self.xd) if i not in self.iaxes]
        rmin, rmax = self._get_radii(self.point, self.xmins0[:, ic], self.xmaxs0[:, ic])
        r_min, r_max = rmin.min(), rmax.max()
        active = np.zeros(self.m, dtype=int)
        for i in range(self.n_steps):
            rt = r_min + (r_max - r_min) * (i / self.n_steps)
            c_radius = (rt >= rmin) & (rt < rmax)
            c_to_active = (active == 0) & c_radius
            c_to_fixed = (active == 1) & (~c_radius)
            if c_to_active.any() or c_to_fixed.any():
                active[c_to_active] = 1
                active[c_to_fixed] = 2
                yield active.copy()
            else:
                yield None


class PointSchedulerRectangularND(_SubspacePointSchedulerRectangularND):
    """schedulers.py:137-143."""

    def __init__(self, all_params, n_steps, point):
        xd = all_params["static"]["decomposition"]["xd"]
        if len(point) != xd:
            raise Exception(f"ERROR: point incorrect shape {np.shape(point)}")
        super().__init__(all_params, n_steps, point, iaxes=[])


class LineSchedulerRectangularND(_SubspacePointSchedulerRectangularND):
    """schedulers.py:145-152."""

    def __init__(self, all_params, n_steps, point, iaxis):
        xd = all_params["static"]["decomposition"]["xd"]
        if xd < 2:
            raise Exception("ERROR: requires nd >=2")
        if len(point) != xd - 1:
            raise Exception(f"ERROR: point incorrect shape {np.shape(point)}")
        super().__init__(all_params, n_steps, point, iaxes=[iaxis])


class PlaneSchedulerRectangularND(_SubspacePointSchedulerRectangularND):
    """schedulers.py:154-161."""

    def __init__(self, all_params, n_steps, point, iaxes):
        xd = all_params["static"]["decomposition"]["xd"]
        if xd < 3:
            raise Exception("ERROR: requires nd >=3")
        if len(point) != xd - 2:
            raise Exception(f"ERROR: point incorrect shape {np.shape(point)}")
        super().__init__(all_params, n_steps, point, iaxes=iaxes)


class SlabSchedulerRectangularND(ActiveScheduler):
    """Time-marching slabs: the subdomain grid is cut along `axis` into slabs
    of `per_slab` subdomain layers (default 1); slab k is active for
    `steps_per_slab` steps, slabs < k are fixed, slabs > k inactive.  Follows
    the Line scheduler's state semantics (schedulers.py:117-135) with a fixed
    step count per slab."""

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
