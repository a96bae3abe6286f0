"""The plain per-op GPU form of the FBPINN train step: the baseline that the
fused engine is measured against.  It is not used by the trainer.

It restates the reference method with stock PyTorch ops, eagerly, op by op:
  1. gather every subdomain parameter leaf per (point, subdomain) pair and
     vmap the per-point plugin functions norm_fn -> network_fn -> unnorm_fn,
     window_fn over the pairs (trainers.py:113-118, 126-158);
  2. nested forward-mode jvp per leaf of the derivative tree
     (trainers.py:63-107, 197-247) with torch.func.jvp;
  3. two index_add segment sums (the POU) back to the points
     (trainers.py:162-170) and constraining_fn;
  4. loss_fn, reverse-mode autograd, and Adam with optax's update rule
     (trainers.py:249-296).
For memory, single-constraint problems may split the points into chunks
whose losses are weighted by their point fraction (exact for the mean-type
losses of the reference problems); gradients are accumulated over chunks
before the Adam update.
"""

import numpy as np
import torch
from torch.func import vmap

from fbpinns_amd.epilogue import get_jmaps, get_ujs
from fbpinns_amd.takes import get_inputs, get_x_batch, split_takes


class PlainFBPINN:
    def __init__(self, problem, decomposition, network, engine, point_chunks=1):
        """Shares the engine's device parameters layout (copied) and points."""
        self.problem, self.decomposition, self.network = problem, decomposition, network
        self.device = engine.device
        ap = engine.all_params
        self.static = ap["static"]
        self.layers = [(W.detach().clone(), b.detach().clone())
                       for W, b in ap["trainable"]["network"]["subdomain"]["layers"]]
        self.prob = {k: v.detach().clone() for k, v in ap["trainable"].get("problem", {}).items()}
        self.lr, self.b1, self.b2, self.eps = engine.lr, engine.b1, engine.b2, engine.eps
        self.mu = [torch.zeros_like(t) for t in self._trainables()]
        self.nu = [torch.zeros_like(t) for t in self._trainables()]
        self.count = 0
        self.point_chunks = int(point_chunks)
        self.engine = engine

    def _trainables(self):
        return [t for W, b in self.layers for t in (W, b)] + list(self.prob.values())

    def set_active(self, active):
        eng = self.engine
        active = np.asarray(active).astype(np.int64)
        ap = {"static": self.static, "trainable": {"network": {"subdomain": {"layers": self.layers}}}}
        tips, cfs, cips, _ = get_x_batch(active, ap, eng.x_global, eng.c_offsets, eng.c_sizes,
                                         self.decomposition)
        x = eng.x_global[torch.as_tensor(tips, dtype=torch.long, device=self.device)]
        takes, all_ims, (active_, *_) = get_inputs(x, active, ap, self.decomposition)
        self.takess = split_takes(takes, cfs, cips, x.shape[0])
        self.all_ims = torch.as_tensor(all_ims, dtype=torch.long, device=self.device)
        self.n_active = int((active_ == 1).sum())
        self.constraints = []
        for ic, ips in enumerate(cips):
            loc = torch.as_tensor(tips[ips] - eng.c_offsets[ic], dtype=torch.long, device=self.device)
            self.constraints.append([v[loc] for v in eng.cglobal[ic]])
        self.jmapss = [get_jmaps(r) for r in eng.required_ujss]
        if self.point_chunks > 1 and len(self.constraints) != 1:
            raise ValueError("point chunking needs a single-constraint problem")

    # ---------------------------------------------------------------- model

    def _model(self, all_params_cut, x_batch, takes):
        """FBPINN_model (trainers.py:126-177) on already-cut parameters."""
        m_take, n_take, p_take, np_take, npou = takes
        x_take = x_batch[n_take]
        sub = {
            "static": {"decomposition": {"subdomain": {
                "params": [p[m_take] for p in all_params_cut["static"]["decomposition"]["subdomain"]["params"]]}}},
            "trainable": {"network": {"subdomain": {
                "layers": [(W[m_take], b[m_take])
                           for W, b in all_params_cut["trainable"]["network"]["subdomain"]["layers"]]}}},
        }
        dec, net = self.decomposition, self.network

        def inner(params, x):
            x_norm = dec.norm_fn(params, x)
            u_raw = net.network_fn(params, x_norm)
            u = dec.unnorm_fn(params, u_raw)
            w = dec.window_fn(params, x)
            return u * w, w, u_raw

        us, ws, _ = vmap(inner)(sub, x_take)
        u = torch.cat([us, ws], dim=1)
        seg = torch.zeros((np_take.shape[0], u.shape[1]), dtype=u.dtype, device=u.device).index_add(0, p_take, u)
        u = seg[:, :-1] / seg[:, -1:]
        u = torch.zeros((x_batch.shape[0], u.shape[1]), dtype=u.dtype, device=u.device).index_add(0, np_take, u)
        u = u / npou
        return self.problem.constraining_fn(all_params_cut, x_batch, u)

    def _loss(self, layers, prob, constraints, takess, weight=1.0):
        ai = self.all_ims
        static_cut = dict(self.static)
        dec = dict(static_cut["decomposition"])
        dec["subdomain"] = {"params": [p[ai] for p in dec["subdomain"]["params"]],
                            "pou": dec["subdomain"]["pou"][ai]}
        static_cut["decomposition"] = dec
        trainable = {"network": {"subdomain": {"layers": [(W[ai], b[ai]) for W, b in layers]}}}
        if prob:
            trainable["problem"] = prob
        ap = {"static": static_cut, "trainable": trainable}
        cs = []
        for takes, jmaps, constraint in zip(takess, self.jmapss, constraints):
            x_batch = constraint[0]
            ujs = get_ujs(x_batch, jmaps, lambda xb, t=takes: (self._model(ap, xb, t), ()))
            cs.append(list(constraint) + ujs)
        return self.problem.loss_fn(ap, cs) * weight

    def _chunks(self):
        """(constraints, takess, weight) per point chunk."""
        if self.point_chunks <= 1:
            t = [tuple(torch.as_tensor(a, dtype=torch.long, device=self.device) if isinstance(a, np.ndarray) else a
                       for a in takes) for takes in self.takess]
            return [(self.constraints, t, 1.0)]
        (m_take, n_take, p_take, np_take, npou), = self.takess
        n = self.constraints[0][0].shape[0]
        bounds = np.linspace(0, n, self.point_chunks + 1).astype(int)
        out = []
        for a, b in zip(bounds[:-1], bounds[1:]):
            sel = (n_take >= a) & (n_take < b)
            segsel = (np_take >= a) & (np_take < b)
            s0 = int(np.nonzero(segsel)[0][0]) if segsel.any() else 0
            dev = lambda v: torch.as_tensor(v, dtype=torch.long, device=self.device)
            tk = (dev(m_take[sel]), dev(n_take[sel] - a), dev(p_take[sel] - s0), dev(np_take[segsel] - a), npou)
            out.append(([[v[a:b] for v in self.constraints[0]]], [tk], (b - a) / n))
        return out

    # ---------------------------------------------------------------- step

    def step(self):
        params = self._trainables()
        leaves = [p.detach().requires_grad_(True) for p in params]
        nl = len(self.layers)
        layers = [(leaves[2 * i], leaves[2 * i + 1]) for i in range(nl)]
        prob = dict(zip(self.prob.keys(), leaves[2 * nl:]))
        grads = [torch.zeros_like(p) for p in params]
        total = 0.0
        for cons, takess, w in self._chunks():
            loss = self._loss(layers, prob, cons, takess, w)
            gs = torch.autograd.grad(loss, leaves, allow_unused=True)
            for g, gi in zip(grads, gs):
                if gi is not None:
                    g.add_(gi)
            total = total + loss.detach()
        self._adam(params, grads)
        return total

    def _adam(self, params, grads):
        """optax.adam update (oracle-independent torch restatement), applied to
        the trained rows only (active subdomains); problem leaves always."""
        self.count += 1
        c = self.count
        bc1 = 1 - self.b1 ** c
        bc2 = 1 - self.b2 ** c
        nl = 2 * len(self.layers)
        rows = self.all_ims[:self.n_active]
        for k, (p, g, m, v) in enumerate(zip(params, grads, self.mu, self.nu)):
            if k < nl:
                gr, mr, vr, pr = g[rows], m[rows], v[rows], p[rows]
            else:
                gr, mr, vr, pr = g, m, v, p
            mr = (1 - self.b1) * gr + self.b1 * mr
            vr = (1 - self.b2) * gr * gr + self.b2 * vr
            upd = (mr / bc1) / (torch.sqrt(vr / bc2) + self.eps)
            pr = pr - self.lr * upd
            if k < nl:
                m[rows], v[rows], p[rows] = mr, vr, pr
            else:
                m.copy_(mr)
                v.copy_(vr)
                p.copy_(pr)

    def model(self, x):
        """Constrained solution u at x with all subdomains (value only)."""
        ap = {"static": self.static, "trainable": {"network": {"subdomain": {"layers": self.layers}}}}
        takes, all_ims, _ = get_inputs(x, np.ones(self.static["decomposition"]["m"], dtype=int), ap,
                                       self.decomposition)
        saved = self.all_ims
        self.all_ims = torch.as_tensor(all_ims, dtype=torch.long, device=self.device)
        ai = self.all_ims
        static_cut = dict(self.static)
        dec = dict(static_cut["decomposition"])
        dec["subdomain"] = {"params": [p[ai] for p in dec["subdomain"]["params"]],
                            "pou": dec["subdomain"]["pou"][ai]}
        static_cut["decomposition"] = dec
        tr = {"network": {"subdomain": {"layers": [(W[ai], b[ai]) for W, b in self.layers]}}}
        if self.prob:
            tr["problem"] = self.prob
        t = tuple(torch.as_tensor(a, dtype=torch.long, device=self.device) if isinstance(a, np.ndarray) else a
                  for a in takes)
        with torch.no_grad():
            u = self._model({"static": static_cut, "trainable": tr}, x, t)
        self.all_ims = saved
        return u
