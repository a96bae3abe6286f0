"""The fused FBPINN train step (replaces FBPINN_update, trainers.py:285-296).

State lives on one GPU in packed form:
  theta [m][P]  every subdomain's FCN parameters (layout in include/fbpx.h);
                all_params["trainable"]["network"]["subdomain"]["layers"]
                holds (W, b) views into it, in the reference's shapes
                (W (m, out, in), b (m, out)).
  mu, nu [m][P] Adam moments; count int32[1] (shared, as optax's).
  pflat [Q]     the problem's trainable leaves (e.g. mu of the inverse
                problem), with their Adam moments.
  dom [m][12]   xmin, xmax, unnorm mean/sd and window flag per subdomain.

One step = fbpx_forward -> fbpx_pou_reduce -> loss epilogue (fbpx_affine_jets
and fbpx_residual_loss, or torch on the per-point jets) -> autograd
(the epilogue's transposes -> fbpx_pou_backward -> fbpx_backward ->
fbpx_reduce_rows) -> fbpx_adam_rows over the trained rows (+ fbpx_adam_flat),
whose last launch also advances the shared step count.  With graph=True the
whole step is captured once per work list into a hipGraph and replayed.
"""

import numpy as np
import torch

from fbpinns_amd import autotune, ops
from fbpinns_amd._lib import FBPX_DOM_STRIDE, FBPX_MAX_XD
from fbpinns_amd.epilogue import Epilogue, schema_masks
from fbpinns_amd.ops import FBPINNJets, FbpxError, NetSpec
from fbpinns_amd.takes import WorkPlan, device_plan, get_x_batch, spatial_order
from fbpinns_amd.util.logger import logger

# compiled jet schemas per input dimension (fbpx_shapes.hpp)
SCHEMAS = {
    1: [(0, 0), (1, 0), (1, 1)],
    2: [(0, 0), (3, 0), (3, 1), (3, 2), (3, 3)],
    3: [(0, 0), (7, 0), (7, 3), (7, 7)],
}


def pick_schema(xd, d1, d2):
    """Smallest compiled schema carrying every requested derivative."""
    best = None
    for a, b in SCHEMAS[xd]:
        if (a & d1) == d1 and (b & d2) == d2:
            j = 1 + bin(a).count("1") + bin(b).count("1")
            if best is None or j < best[0]:
                best = (j, a, b)
    if best is None:
        raise FbpxError(f"no compiled jet schema carries d1mask={d1} d2mask={d2} for xd={xd}")
    return best[1], best[2]


def tree_to(t, device):
    if isinstance(t, dict):
        return {k: tree_to(v, device) for k, v in t.items()}
    if isinstance(t, list):
        return [tree_to(v, device) for v in t]
    if isinstance(t, tuple):
        return tuple(tree_to(v, device) for v in t)
    if isinstance(t, torch.Tensor):
        return t.to(device)
    return t


def network_shape(layers):
    """(xd, h, nh) of a stacked FCN, or raise if it is not [xd, h x nh, 1]."""
    sizes = [layers[0][0].shape[2]] + [W.shape[1] for W, _ in layers]
    xd, hs, ud = sizes[0], sizes[1:-1], sizes[-1]
    if ud != 1:
        raise FbpxError(f"the fused engine computes scalar solutions; network output size is {ud}")
    if len(set(hs)) != 1 or len(hs) not in (1, 2):
        raise FbpxError(f"layer sizes {sizes} not compiled: need [xd, h] or [xd, h, h] hidden layers")
    if xd < 1 or xd > FBPX_MAX_XD:
        raise FbpxError(f"input dimension {xd} not compiled (1..{FBPX_MAX_XD})")
    return xd, hs[0], len(hs)


class FusedEngine:
    """Owns the packed device state and runs train steps for one problem.

    graph: capture the whole step in a hipGraph; autotune: time candidate
    launch shapes per work list; fused_loss: run a problem's residual
    description through the fused residual kernels (else its loss_fn);
    sort_points: keep each constraint's points in Z-order when the problem
    has a residual description (its loss is then order-independent)."""

    def __init__(self, problem, decomposition, network, all_params, optimiser_kwargs=None, device="cuda:0",
                 graph=True, autotune=True, fused_loss=True, sort_points=True, fold_adam=True):
        from fbpinns_amd.decompositions import RectangularDecompositionND
        from fbpinns_amd.networks import FCN
        if not torch.cuda.is_available():
            raise FbpxError("the fused engine needs a GPU (torch.cuda.is_available() is False)")
        if not (isinstance(decomposition, type) and issubclass(decomposition, RectangularDecompositionND)):
            raise FbpxError("the fused engine accelerates RectangularDecompositionND (and its multilevel form)")
        if not (isinstance(network, type) and issubclass(network, FCN)):
            raise FbpxError("the fused engine accelerates the tanh FCN network")
        self.problem, self.decomposition, self.network = problem, decomposition, network
        self.device = torch.device(device)
        self.graph_enabled = bool(graph)
        self.autotune_enabled = bool(autotune)
        self.fused_loss = bool(fused_loss)
        self.sort_points = bool(sort_points)
        self.fold_adam = bool(fold_adam)
        ok = dict(optimiser_kwargs or {})
        self.lr = float(ok.pop("learning_rate", 1e-3))
        self.b1 = float(ok.pop("b1", 0.9))
        self.b2 = float(ok.pop("b2", 0.999))
        self.eps = float(ok.pop("eps", 1e-8))
        if float(ok.pop("eps_root", 0.0)) != 0.0 or ok.pop("nesterov", False) or ok:
            raise FbpxError(f"unsupported Adam options {ok}")

        ap = tree_to(all_params, self.device)
        self.all_params = ap
        dec = ap["static"]["decomposition"]
        self.m, self.xd = int(dec["m"]), int(dec["xd"])
        layers = ap["trainable"]["network"]["subdomain"]["layers"]
        xd, h, nh = network_shape(layers)
        if xd != self.xd:
            raise FbpxError("network input size != decomposition dimension")
        self.net0 = NetSpec(xd, h, nh)
        self.P = self.net0.P
        self.theta = torch.empty((self.m, self.P), dtype=torch.float32, device=self.device)
        views = []
        off = 0
        for W, b in layers:
            o, i = W.shape[1], W.shape[2]
            Wv = self.theta[:, off:off + o * i].view(self.m, o, i)
            Wv.copy_(W)
            off += o * i
            bv = self.theta[:, off:off + o]
            bv.copy_(b)
            off += o
            views.append((Wv, bv))
        assert off == self.P
        ap["trainable"]["network"]["subdomain"]["layers"] = views
        self.mu = torch.zeros_like(self.theta)
        self.nu = torch.zeros_like(self.theta)
        self.count = torch.zeros(1, dtype=torch.int32, device=self.device)
        self.count_done = torch.zeros(1, dtype=torch.int32, device=self.device)   # 0 between launches
        self.loss_seed = torch.ones((), dtype=torch.float32, device=self.device)

        # problem trainable leaves -> one flat vector with views
        prob = ap["trainable"].get("problem", {})
        self.pkeys = list(prob.keys())
        self.pshapes = [tuple(torch.as_tensor(prob[k]).shape) for k in self.pkeys]
        Q = sum(int(np.prod(s)) for s in self.pshapes)
        self.pflat = torch.zeros(max(Q, 1), dtype=torch.float32, device=self.device)
        self.Q = Q
        off = 0
        for k, s in zip(self.pkeys, self.pshapes):
            n = int(np.prod(s))
            self.pflat[off:off + n].copy_(torch.as_tensor(prob[k], dtype=torch.float32).reshape(-1))
            prob[k] = self.pflat[off:off + n].view(s)
            off += n
        self.pmu = torch.zeros_like(self.pflat)
        self.pnu = torch.zeros_like(self.pflat)

        # per-subdomain static rows
        p = dec["subdomain"]["params"]
        dom = torch.zeros((self.m, FBPX_DOM_STRIDE), dtype=torch.float32, device=self.device)
        dom[:, 0:self.xd] = p[0]
        dom[:, 4:4 + self.xd] = p[1]
        dom[:, 8] = p[5][:, 0]
        dom[:, 9] = p[5][:, 1]
        dom[:, 10] = p[4][:, 0]
        self.dom = dom.contiguous()
        self.plan = None
        self._graph = None
        self._test_cache = {}

    # -------------------------------------------------------------- params

    def _params_for_step(self):
        """all_params whose trainable problem leaves are differentiable views
        of a leaf that shares pflat's storage."""
        ap = self.all_params
        if not self.Q:
            return ap, None
        leaf = self.pflat.detach().requires_grad_(True)
        prob = {}
        off = 0
        for k, s in zip(self.pkeys, self.pshapes):
            n = int(np.prod(s))
            # a leaf spanning pflat is a view of it: its backward is a reshape,
            # where a slice's backward fills a zero tensor and copies into it
            prob[k] = leaf.view(s) if n == leaf.numel() else leaf[off:off + n].view(s)
            off += n
        tr = dict(ap["trainable"])
        tr["problem"] = prob
        return {"static": ap["static"], "trainable": tr}, leaf

    # -------------------------------------------------------------- optimiser state

    def _tree_of(self, rows, flat):
        """A tree mirroring all_params["trainable"] whose leaves are views of
        the packed rows [m][P] (network layers) and flat [Q] (problem leaves)."""
        layers, off = [], 0
        for W, b in self.all_params["trainable"]["network"]["subdomain"]["layers"]:
            o, i = W.shape[1], W.shape[2]
            layers.append((rows[:, off:off + o * i].view(self.m, o, i), rows[:, off + o * i:off + o * i + o]))
            off += o * i + o
        prob, off = {}, 0
        for k, shp in zip(self.pkeys, self.pshapes):
            n = int(np.prod(shp))
            prob[k] = flat[off:off + n].view(shp)
            off += n
        tree = {}
        for key in self.all_params["trainable"]:
            if key == "network":
                tree[key] = {"subdomain": {"layers": layers}}
            elif key == "problem":
                tree[key] = prob
        return tree

    def opt_state(self):
        """optax.adam's state as (count, mu, nu): the shared int32 step count
        and the first / second moments as trees mirroring
        all_params["trainable"] (trainers.py:430-434)."""
        return self.count.reshape(()), self._tree_of(self.mu, self.pmu), self._tree_of(self.nu, self.pnu)

    def load_opt_state(self, count, mu, nu):
        """Set the Adam state from (count, mu, nu) in the layout opt_state()
        returns (e.g. read back from a checkpoint)."""
        for dst_rows, dst_flat, tree in ((self.mu, self.pmu, mu), (self.nu, self.pnu, nu)):
            view = self._tree_of(dst_rows, dst_flat)
            for (W, b), (sW, sb) in zip(view["network"]["subdomain"]["layers"], tree["network"]["subdomain"]["layers"]):
                W.copy_(torch.as_tensor(np.asarray(sW)))
                b.copy_(torch.as_tensor(np.asarray(sb)))
            for k in self.pkeys:
                view["problem"][k].copy_(torch.as_tensor(np.asarray(tree["problem"][k]), dtype=torch.float32))
        self.count.fill_(int(np.asarray(count)))

    # -------------------------------------------------------------- set-up

    def set_constraints(self, constraints_global, required_ujss):
        """constraints_global: [[x_c, *values_c], ...] (CPU or GPU tensors)."""
        self.cglobal = [[tree_to(c_, self.device) for c_ in c] for c in constraints_global]
        if self.sort_points and self.problem.residuals(self.all_params) is not None:
            # a loss stated as means of pointwise residuals does not depend on
            # the order of a constraint's points (up to the summation order):
            # order them along a space-filling curve, for locality in the
            # kernels' gathers
            for c in self.cglobal:
                if not all(isinstance(v, torch.Tensor) and v.shape[:1] == c[0].shape[:1] for v in c):
                    raise FbpxError("constraint values must be per-point tensors")
                perm = torch.as_tensor(spatial_order(c[0].detach().cpu().numpy()), dtype=torch.long,
                                       device=self.device)
                c[:] = [v[perm].contiguous() for v in c]
        self.required_ujss = [tuple((iu, tuple(ixs)) for iu, ixs in r) for r in required_ujss]
        self.x_global = torch.cat([c[0] for c in self.cglobal]).to(torch.float32).contiguous()
        sizes = [c[0].shape[0] for c in self.cglobal]
        self.c_sizes = sizes
        self.c_offsets = list(np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(int))
        d1, d2 = schema_masks(self.required_ujss, self.xd)
        d1, d2 = pick_schema(self.xd, d1, d2)
        self.net = self.net0.with_schema(d1, d2)
        if not self.net.supported():
            raise FbpxError(f"no kernels for {self.net}")

    def set_active(self, active, chunk=None, launch=None):
        """New scheduler state: rebuild takes, work list, epilogue and graph
        (trainers.py:509-575 and the recompile at :647-653).  launch:
        (forward, backward) launch shapes, each ("chunks", C), ("bins", k) or,
        for one-hidden-layer networks, ("waves", C) (see autotune); chunk=C
        is (("chunks", C), ("chunks", C)).  Default:
        autotuned, or autotune.default_spec with autotune off."""
        active = np.asarray(active).astype(np.int64)
        training_ips, cfs, cips, dstat = get_x_batch(active, self.all_params, self.x_global, self.c_offsets,
                                                     self.c_sizes, self.decomposition)
        tips = torch.as_tensor(training_ips, dtype=torch.long, device=self.device)
        x = self.x_global[tips].contiguous()
        cranges, cvalues = [], []
        off = 0
        for ic, ips in enumerate(cips):
            n_c = len(ips)
            loc = torch.as_tensor(training_ips[ips] - self.c_offsets[ic], dtype=torch.long, device=self.device)
            cvalues.append([x[off:off + n_c]] + [v[loc] for v in self.cglobal[ic][1:]])
            cranges.append((off, n_c))
            off += n_c
        # the takes and the default small-list work list in one device pass
        small = autotune.DEFAULT_SMALL
        plan0, takes, all_ims, active_ = device_plan(x, active, self.all_params, chunk=small[1])
        n_active = plan0.n_active
        n_pairs = plan0.wl.n_pairs
        slots = []

        def build(fspec, bspec):
            waves = (fspec[0] == "waves", bspec[0] == "waves")
            if any(waves):   # the chunk work list, run one wave per bin
                if self.net.nh != 1:
                    raise FbpxError(f"one wave per bin runs one-hidden-layer networks only, not {self.net}")
                p = build(*[("chunks", c[1]) if c[0] == "waves" else c for c in (fspec, bspec)])
                return WorkPlan(p.wl.with_wave_bins(*waves), p.chunk_train, p.rows, p.row_chunk_start,
                                p.n_active, p.all_ims)
            if fspec == bspec == small:
                return plan0
            caps = [c[1] for c in (fspec, bspec) if c[0] == "chunks"]
            fb = bb = None
            if "bins" in (fspec[0], bspec[0]):
                if not slots:   # resident workgroups of each kernel on the whole GPU
                    fo, bo = self.net.occupancy()
                    n_cu = torch.cuda.get_device_properties(self.device).multi_processor_count
                    slots.extend([max(1, fo) * n_cu, max(1, bo) * n_cu])
                fb = fspec[1] * slots[0] if fspec[0] == "bins" else None
                bb = bspec[1] * slots[1] if bspec[0] == "bins" else None
            return device_plan(x, active, self.all_params, chunk=min(caps) if caps else None, fbins=fb,
                               bbins=bb)[0]
        if chunk is not None:
            launch = (("chunks", int(chunk)), ("chunks", int(chunk)))
        if launch is not None:
            self.launch, plan = (tuple(launch[0]), tuple(launch[1])), build(tuple(launch[0]), tuple(launch[1]))
        else:
            n_cu = torch.cuda.get_device_properties(self.device).multi_processor_count
            if self.autotune_enabled:
                self.launch, plan = autotune.tune_plan(build, self.net, self.theta, self.dom, x, self.m, n_pairs,
                                                       n_chunks=plan0.wl.n_chunks, n_cu=n_cu)
            else:
                spec = autotune.default_spec(n_pairs, self.net.nh, plan0.wl.n_chunks, n_cu)
                self.launch, plan = (spec, spec), build(spec, spec)
        wsum = torch.empty((max(plan.wl.n_segs, 1), self.net.J), dtype=torch.float32, device=self.device)
        scratch = torch.empty((max(plan.wl.n_pairs, 1), self.net.J), dtype=torch.float32, device=self.device)
        ops.pou_prepare(self.net.as_list(), self.dom, x, plan.wl.tensor_list(), plan.wl.size_list(), scratch, wsum)
        del scratch
        self.plan = ops.Plan(self.net, self.dom, x, plan.wl, wsum, plan.chunk_train, plan.rows,
                             plan.row_chunk_start, self.m)
        self.plan.keep_gradient_buffer(self.net.P)
        # the step's Adam inside the backward's launch (fbpx_backward_adam):
        # one hidden layer, no trainable problem leaves, and every trained
        # row exactly one chunk (so each chunk's row is its whole gradient)
        rcs = plan.row_chunk_start.cpu()
        if (self.fold_adam and self.net.nh == 1 and not self.Q and len(rcs) > 1
                and bool(((rcs[1:] - rcs[:-1]) == 1).all())):
            self.plan.adam_fold = (self.theta, self.mu, self.nu, self.count, self.count_done, self.lr, self.b1,
                                   self.b2, self.eps)
        self.workplan = plan
        self.active = active_
        self.x = x
        self.takes = takes
        self.epilogue = Epilogue(self.problem, self.all_params, x, cranges, cvalues, self.required_ujss, self.net,
                                 fused_loss=self.fused_loss)
        self.n_active = n_active
        logger.info(f"work list: {plan.wl.n_points} points, {plan.wl.n_pairs} pairs, {plan.wl.n_chunks} chunks, "
                    f"fwd {autotune.name(self.launch[0])} ({plan.wl.workgroups()[0]} workgroups), "
                    f"bwd {autotune.name(self.launch[1])} ({plan.wl.workgroups()[1]} workgroups), "
                    f"{n_active} trained / {len(all_ims)} used subdomains, jets J={self.net.J}, "
                    f"epilogue={'affine' if self.epilogue.affine is not None else 'generic'}"
                    f"{' + fused residual loss' if self.epilogue.fused is not None else ''}, d={dstat:.2f}")
        self._graph = None
        if self.graph_enabled:
            self._capture()
        return dstat

    # -------------------------------------------------------------- step

    def loss_and_grad(self):
        """Loss and gradients at the current parameters without an update:
        (loss, d loss/d theta [m][P], d loss/d pflat [Q] or None)."""
        plan = self.plan
        plan.defer_rows = False
        th = self.theta.detach().requires_grad_(True)
        params, pleaf = self._params_for_step()
        loss = self.epilogue.jets_loss(th, plan, params)
        leaves = [th] + ([pleaf] if pleaf is not None else [])
        grads = torch.autograd.grad(loss, leaves, allow_unused=True)
        g_p = None
        if pleaf is not None:
            g_p = grads[1] if grads[1] is not None else torch.zeros_like(self.pflat)
        return loss.detach(), grads[0].clone(), g_p

    def _step_body(self):
        plan = self.plan
        # the backward leaves the per-chunk gradient rows in plan.partial;
        # adam_rows_partial sums and applies them in one launch (or, with
        # plan.adam_fold, the backward applies Adam itself)
        plan.defer_rows = True
        plan.adam_folded = False
        th = self.theta.detach().requires_grad_(True)
        params, pleaf = self._params_for_step()
        loss = self.epilogue.jets_loss(th, plan, params)
        # the preallocated seed d loss / d loss = 1 keeps a fill kernel out of the step
        leaves = [th] + ([pleaf] if pleaf is not None else [])
        grads = torch.autograd.grad(loss, leaves, grad_outputs=(self.loss_seed,), allow_unused=True)
        g_p = grads[1] if pleaf is not None else None
        # the step's last Adam launch also advances the shared count
        if not plan.adam_folded:
            ops.adam_rows_partial(self.theta, self.mu, self.nu, plan.partial, plan.rows, plan.row_chunk_start,
                                  self.count, self.count_done, not self.Q, self.lr, self.b1, self.b2, self.eps)
        if self.Q:
            if g_p is None:
                g_p = torch.zeros_like(self.pflat)
            ops.adam_flat(self.pflat, self.pmu, self.pnu, g_p.contiguous(), self.count, self.count_done, True,
                          self.lr, self.b1, self.b2, self.eps)
        return loss.detach()

    def _state(self):
        return [self.theta, self.mu, self.nu, self.pflat, self.pmu, self.pnu, self.count]

    def _capture(self):
        saved = [t.clone() for t in self._state()]
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(2):
                self._step_body()
        torch.cuda.current_stream().wait_stream(s)
        for t, v in zip(self._state(), saved):
            t.copy_(v)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._graph_loss = self._step_body()
        self._graph = g
        torch.cuda.synchronize()

    def step(self):
        """One train step; returns the loss as a device scalar."""
        if self.plan is None:
            raise FbpxError("set_active() must be called before step()")
        if self._graph is not None:
            self._graph.replay()
            return self._graph_loss
        return self._step_body()

    # -------------------------------------------------------------- inference

    def make_plan(self, x, active=None, schema=(0, 0)):
        """Plan (work list + buffers) for points x with every subdomain (or
        `active`) used and every used subdomain trained."""
        x = x.to(self.device, torch.float32).contiguous()
        active = np.ones(self.m, dtype=int) if active is None else np.asarray(active)
        plan, takes, _, _ = device_plan(x, active, self.all_params)
        net = self.net0.with_schema(*schema)
        wsum = torch.empty((max(plan.wl.n_segs, 1), net.J), dtype=torch.float32, device=self.device)
        scratch = torch.empty((max(plan.wl.n_pairs, 1), net.J), dtype=torch.float32, device=self.device)
        ops.pou_prepare(net.as_list(), self.dom, x, plan.wl.tensor_list(), plan.wl.size_list(), scratch, wsum)
        return ops.Plan(net, self.dom, x, plan.wl, wsum, plan.chunk_train, plan.rows, plan.row_chunk_start,
                        self.m), takes

    def model_jets(self, x, active=None, schema=(0, 0)):
        """Jets of u~ (before constraining_fn) at points x; returns (J, n)."""
        x = x.to(self.device, torch.float32).contiguous()
        key = (x.data_ptr(), x.shape[0], None if active is None else tuple(np.asarray(active).tolist()), schema)
        if key not in self._test_cache:
            self._test_cache = {key: (x, self.make_plan(x, active, schema)[0])}
        x, plan = self._test_cache[key]
        with torch.no_grad():
            return FBPINNJets.apply(self.theta, plan)

    def model(self, x, active=None):
        """Constrained FBPINN solution u (n, 1) at x (FBPINN_model, trainers.py:126-177)."""
        x = x.to(self.device, torch.float32).contiguous()
        ujet = self.model_jets(x, active)
        params, _ = self._params_for_step()
        with torch.no_grad():
            return self.problem.constraining_fn(params, x, ujet[0][:, None])
