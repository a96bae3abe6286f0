"""Loss epilogue: from the kernels' per-point jets of the POU-combined network
output (u~ and its derivatives) to the problem's loss.

The reference applies `constraining_fn` inside the model and differentiates
the whole model with one nested forward-mode jvp per leaf of the derivative
tree (trainers.py:63-107, 174, 197-247).  Here the expensive part (the
subdomain networks and the POU sum) arrives as exact jets from the HIP
kernels, and only the cheap constraining function is differentiated:

* affine path (constraining_fn(x, u) = A(x) + B(x) u with A, B independent of
  trainable parameters, detected numerically at set-up): the jets of A and B
  are computed once, in float64, with the same nested jvp, and each step
  applies the Leibniz rule in one kernel per constraint (fbpx_affine_jets and
  its transpose, ops.AffineJets);
* generic path (anything else): u~ is replaced near each point by its exact
  second-order Taylor polynomial in the requested axes, and the reference's
  nested jvp tree is evaluated on constraining_fn(x, poly(x)), which gives
  the same derivatives for every request of order <= 2 with repeated indices.

Then loss_fn(all_params, constraints_) is evaluated exactly as the reference
calls it (trainers.py:258-267), or, when the problem describes its loss as
weighted means of squared pointwise residuals (fbpinns_amd.residuals) and the
description agrees with loss_fn at set-up, one fused kernel pair per residual
(fbpx_residual_loss, ops.ResidualLoss) evaluates it on the GPU.
"""

import torch
from torch.func import jvp

from fbpinns_amd import _lib, ops
from fbpinns_amd._lib import FBPX_MAX_REQ
from fbpinns_amd.ops import AffineJets, FBPINNJets, FbpxError, GatherRows, ResidualLoss
from fbpinns_amd.residuals import F, J, P, V, check_spec, spec_loss
from fbpinns_amd.util.logger import logger


# ------------------------------------------------------------------ derivative tree

def get_jmaps(required_ujs):
    """Derivative tree of required_ujs (trainers.py:63-107)."""
    tree = {}
    for iu, ixs in required_ujs:
        t = tree
        for ix in ixs:
            t = t.setdefault(ix, {})

    def nodes_of(t, n=(), ks=()):
        ni = len(n)
        for k in t:
            ks_ = ks + (k,)
            if t[k]:
                n += (((ni, k), ks_, 0),)
                n = nodes_of(t[k], n, ks_)
            else:
                n += (((ni, k), ks_, 1),)
        return n

    nodes = nodes_of(tree)
    leaves = tuple((i + 1, node[1]) for i, node in enumerate(nodes) if node[2]) or ((0, ()),)
    jac_is = tuple(([leaf[1][:len(ixs)] for leaf in leaves].index(ixs), len(ixs), iu) for iu, ixs in required_ujs)
    return nodes, leaves, jac_is


def get_ujs(x_batch, jmaps, u):
    """Evaluate the nested-jvp tree (trainers.py:213-247); u(x) -> ((n, ud), ())."""
    nodes, leaves, jac_is = jmaps
    n, xd = x_batch.shape
    eye = torch.eye(xd, dtype=x_batch.dtype, device=x_batch.device)

    def jacfwd(f, v):
        def jacfun(x):
            y, j, aux = jvp(f, (x,), (v,), has_aux=True)
            return j, aux + (y,)
        return jacfun

    fs = [u]
    for (ni, ix), _, _ in nodes:
        fs.append(jacfwd(fs[ni], eye[ix].expand(n, xd)))
    jacs = []
    for ie, _ in leaves:
        fin, jac = fs[ie](x_batch)
        jacs.append(jac + (fin,))
    return [jacs[il][io][:, iu:iu + 1] for il, io, iu in jac_is]


def schema_masks(required_ujss, xd):
    """(d1mask, d2mask) of u~ jets needed by every constraint's requests."""
    d1 = d2 = 0
    for req in required_ujss:
        for iu, ixs in req:
            if iu != 0:
                raise FbpxError("the fused engine computes scalar solutions (ud == 1); "
                                f"request {(iu, ixs)} asks for component {iu}")
            ixs = tuple(ixs)
            if any(i < 0 or i >= xd for i in ixs):
                raise FbpxError(f"derivative index out of range in {(iu, ixs)}")
            if len(ixs) == 1:
                d1 |= 1 << ixs[0]
            elif len(ixs) == 2:
                if ixs[0] != ixs[1]:
                    raise FbpxError(f"mixed second derivative {ixs} is not carried by the jet kernels "
                                    "(diagonal second derivatives only)")
                d1 |= 1 << ixs[0]
                d2 |= 1 << ixs[0]
            elif len(ixs) > 2:
                raise FbpxError(f"derivative order {len(ixs)} > 2 is not carried by the jet kernels")
    return d1, d2


def _col_index(net):
    """(kind, dim) -> row of the jet array."""
    idx = {("v", None): 0}
    for i in range(net.xd):
        if (net.d1mask >> i) & 1:
            idx[("d1", i)] = net.col1(i)
        if (net.d2mask >> i) & 1:
            idx[("d2", i)] = net.col2(i)
    return idx


class Epilogue:
    """Maps kernel jets to the problem loss for one set of training points.

    x          (n, xd) float32 training points (all constraints, concatenated)
    cranges    [(offset, size)] of each constraint inside x
    cvalues    per constraint, the list [x_c, *values_c] passed to loss_fn
    """

    def __init__(self, problem, all_params, x, cranges, cvalues, required_ujss, net, fused_loss=True):
        self.problem, self.net = problem, net
        self.x, self.cranges, self.cvalues = x, cranges, cvalues
        self.required_ujss = [tuple((iu, tuple(ixs)) for iu, ixs in r) for r in required_ujss]
        self.jmapss = [get_jmaps(r) for r in self.required_ujss]
        self.col = _col_index(net)
        self.affine = self._affine_precompute(all_params)
        self.fused = self._fused_precompute(all_params) if fused_loss else None
        self.combined = self._combined_setup()
        # d loss / d (constrained jets): columns no residual reads stay 0
        self.gC = (torch.zeros((net.J, x.shape[0]), dtype=torch.float32, device=x.device)
                   if self.combined is not None else None)
        # the fused loss's forward that last wrote gC (or the plan's gseg) at
        # unit upstream gradient (its token), so that only its own backward
        # reads it there
        self._gC_owner = None
        self._grad_wanted = False
        # fbpx_pou_loss takes every residual in one launch when they fit it
        self.pou_loss_fits = self.combined is not None and len(self.fused) <= _lib.FBPX_POU_LOSS_RES and all(
            len(keys) <= _lib.FBPX_POU_LOSS_COLS for _, _, keys, *_ in self.fused)
        self._pou_loss_done = (torch.zeros(1, dtype=torch.int32, device=x.device) if self.pou_loss_fits else None)
        # fbpx_residual_losses runs two or more of them in one launch (the
        # same limits; one residual keeps its own launch)
        self.residual_set_fits = self.pou_loss_fits and len(self.fused) >= 2

    # -------------------------------------------------------------- set-up

    def _affine_precompute(self, all_params):
        C = self.problem.constraining_fn
        x64 = self.x.detach().double()
        n = x64.shape[0]
        if n == 0:
            return None
        # independence from trainable parameters
        probe = _with_grad_trainable(all_params)
        u0 = torch.zeros((n, 1), dtype=torch.float64, device=x64.device)
        try:
            out = C(probe, x64, u0)
        except Exception:  # noqa: BLE001 - any failure selects the generic path
            return None
        if not isinstance(out, torch.Tensor) or out.requires_grad or out.shape != (n, 1):
            return None
        # affinity in u
        g = torch.Generator(device="cpu").manual_seed(1234)
        r = (torch.rand((n, 1), generator=g, dtype=torch.float64) * 4 - 2).to(x64.device)
        with torch.no_grad():
            A = C(all_params, x64, u0)
            B = C(all_params, x64, u0 + 1) - A
            Cr = C(all_params, x64, r)
        scale = 1.0 + A.abs().max() + B.abs().max()
        if not torch.isfinite(scale) or (Cr - (A + B * r)).abs().max() > 1e-9 * scale:
            return None
        # jets of A and B in the schema's columns (float64, nested jvp)
        req = [(0, ())]
        for i in range(self.net.xd):
            if (self.net.d1mask >> i) & 1:
                req.append((0, (i,)))
            if (self.net.d2mask >> i) & 1:
                req.append((0, (i, i)))
        jm = get_jmaps(tuple(req))

        def fA(xb):
            return C(all_params, xb, torch.zeros((xb.shape[0], 1), dtype=xb.dtype, device=xb.device)), ()

        def fB(xb):
            z = torch.zeros((xb.shape[0], 1), dtype=xb.dtype, device=xb.device)
            return C(all_params, xb, z + 1) - C(all_params, xb, z), ()

        dt = self.x.dtype
        Aj = torch.cat(get_ujs(x64, jm, fA), 1).T.contiguous().to(dt)   # (J, n) in req order
        Bj = torch.cat(get_ujs(x64, jm, fB), 1).T.contiguous().to(dt)
        # req order equals the jet column order (value, then per axis d1/d2 interleaved) -> remap
        order = []
        for iu, ixs in req:
            order.append(self.col[("v", None)] if len(ixs) == 0 else
                         self.col[("d1", ixs[0])] if len(ixs) == 1 else self.col[("d2", ixs[0])])
        A_cols = torch.empty_like(Aj)
        B_cols = torch.empty_like(Bj)
        A_cols[order] = Aj
        B_cols[order] = Bj
        return A_cols, B_cols

    def _fused_precompute(self, all_params):
        """Residual description of loss_fn (problem.residuals), checked against
        loss_fn in float64, with the per-point data columns and static
        coefficients it needs each step; None selects loss_fn itself."""
        fn = getattr(self.problem, "residuals", None)
        if fn is None or not self.x.is_cuda:
            return None
        n_req = [len(r) for r in self.required_ujss]
        n_values = [len(c) - 1 for c in self.cvalues]
        try:
            spec = fn(all_params)
            if spec is None:
                return None
            check_spec(spec, len(self.cranges), n_req, n_values)
            if any(self.cranges[res.constraint][1] == 0 for res in spec):
                return None
            if not self._fused_agrees(spec, all_params, n_values):
                logger.info("fused residual loss: description disagrees with loss_fn, using loss_fn")
                return None
            return self._fused_plans(spec, all_params)
        except Exception as e:  # noqa: BLE001 - any failure selects loss_fn
            logger.warning(f"fused residual loss not used ({type(e).__name__}: {e}); running loss_fn")
            return None

    def _fused_plans(self, spec, all_params):
        """Per residual: (constraint, weight, column keys, per-point data,
        factor columns fa / fb, static coefficients, trainable coefficients,
        the zeroed workgroup counter of its single-launch kernels)."""
        plans = []
        for res in spec:
            ic = res.constraint
            x_c = self.cvalues[ic][0]
            n_c = self.cranges[ic][1]
            keys, data, fa, fb = [], {}, [], []

            def slot(f):
                if f is None:
                    return -1
                key = ("J", f.i) if isinstance(f, J) else ("V", f.j) if isinstance(f, V) else \
                    ("P", id(f)) if isinstance(f, P) else ("F", id(f))
                if key not in keys:
                    keys.append(key)
                    vals = [v.double() if v.is_floating_point() else v for v in self.cvalues[ic][1:]]
                    if isinstance(f, J):
                        pass
                    elif isinstance(f, V):
                        data[key] = self.cvalues[ic][1 + f.j].to(torch.float32).reshape(n_c).contiguous()
                    elif isinstance(f, F):
                        with torch.no_grad():
                            v = f.fn(all_params, x_c.double(), *vals)
                        data[key] = torch.as_tensor(v).to(torch.float64).expand(n_c, 1).reshape(n_c) \
                            .to(torch.float32).contiguous()
                    else:
                        # static index: points sorted by index once, for the
                        # deterministic transpose of the gather
                        idx = self.cvalues[ic][1 + f.j].reshape(n_c).long()
                        n_vec = f.fn(all_params).numel()
                        if idx.min() < 0 or idx.max() >= n_vec:
                            raise ValueError(f"gather index outside [0, {n_vec})")
                        perm = torch.argsort(idx, stable=True).to(torch.int32)
                        counts = torch.bincount(idx, minlength=n_vec)
                        seg = torch.zeros(n_vec + 1, dtype=torch.int64, device=idx.device)
                        seg[1:] = torch.cumsum(counts, 0)
                        data[key] = (f.fn, idx.to(torch.int32).contiguous(), perm.contiguous(),
                                     seg.to(torch.int32).contiguous())
                return keys.index(key)
            for _, a, b in res.terms:
                fa.append(slot(a))
                fb.append(slot(b))
            if len(keys) > FBPX_MAX_REQ:
                return None
            static = torch.tensor([0.0 if callable(c) else float(c) for c, _, _ in res.terms], dtype=torch.float32,
                                  device=self.x.device)
            dyn = [(t, c) for t, (c, _, _) in enumerate(res.terms) if callable(c)]
            done = torch.zeros(1, dtype=torch.int32, device=self.x.device)   # single-launch counter
            plans.append((ic, float(res.weight), keys, data, fa, fb, static, dyn, done))
        return plans

    def _fused_agrees(self, spec, all_params, n_values):
        """loss_fn and the residual description at random derivatives, in
        float64: values and gradients (derivatives, trainable parameters)."""
        probe = _with_grad_trainable(all_params)
        leaves = _trainable_leaves(probe)
        g = torch.Generator(device="cpu").manual_seed(4321)
        cons, jets = [], []
        for ic, (off, n_c) in enumerate(self.cranges):
            vals = [v.double() if isinstance(v, torch.Tensor) and v.is_floating_point() else v
                    for v in self.cvalues[ic]]
            js = [torch.randn((n_c, 1), generator=g, dtype=torch.float64).to(self.x.device).requires_grad_(True)
                  for _ in self.required_ujss[ic]]
            cons.append(vals + js)
            jets += js
        a = self.problem.loss_fn(probe, cons)
        b = spec_loss(spec, probe, cons, n_values)
        wrt = jets + leaves
        ga = torch.autograd.grad(a, wrt, allow_unused=True)
        gb = torch.autograd.grad(b, wrt, allow_unused=True)
        a, b = float(a.detach()), float(b.detach())
        if not abs(a - b) <= 1e-9 * max(abs(a), 1e-300):
            return False
        for u, v, w in zip(ga, gb, wrt):
            u = torch.zeros_like(w) if u is None else u
            v = torch.zeros_like(w) if v is None else v
            if (u - v).abs().max() > 1e-9 * (1.0 + u.abs().max()):
                return False
        return True

    @staticmethod
    def _coef(static, dyn, all_params):
        if not dyn:
            return static
        parts = list(static.unbind())
        for t, c in dyn:
            parts[t] = torch.as_tensor(c(all_params)).reshape(()).to(torch.float32)
        return torch.stack(parts)

    def _combined_setup(self):
        """Jet column of every derivative key when the jets kernel, the affine
        constraining function and the residuals can run as one autograd node
        (_JetsResidualLoss): affine constraint, fused residuals, every
        (constraint, jet column) read by one residual only (its gradient is
        written, not accumulated).  Returns the per-residual column lists
        (jet column, or None for data and gathered columns) or None."""
        if self.affine is None or self.fused is None:
            return None
        cols, seen = [], set()
        for ic, weight, keys, data, fa, fb, static, dyn, done in self.fused:
            cs = []
            for k in keys:
                if k[0] == "J":
                    iu, ixs = self.required_ujss[ic][k[1]]
                    c = (self.col[("v", None)] if len(ixs) == 0 else
                         self.col[("d1", ixs[0])] if len(ixs) == 1 else self.col[("d2", ixs[0])])
                    if (ic, c) in seen:
                        return None
                    seen.add((ic, c))
                    cs.append(c)
                else:
                    cs.append(None)
            cols.append(cs)
        return cols

    def jets_loss(self, theta, plan, all_params):
        """The loss from the subdomain rows theta on the plan's points
        (FBPINN_loss, trainers.py:249-267): one autograd node for jets,
        constraint and residuals where _combined_setup allows, else the jets
        (ops.FBPINNJets) followed by Epilogue.loss."""
        if self.combined is not None:
            coefs = [self._coef(static, dyn, all_params) for _, _, _, _, _, _, static, dyn, _ in self.fused]
            # the trainable vectors of the gathered columns, in key order
            pvecs = [torch.as_tensor(data[k][0](all_params)).reshape(-1).to(torch.float32)
                     for _, _, keys, data, *_ in self.fused for k in keys if k[0] == "P"]
            # inside the node's forward grad mode is off; whether a backward
            # can follow is decided here
            self._grad_wanted = torch.is_grad_enabled() and theta.requires_grad
            return _JetsResidualLoss.apply(theta, plan, self, len(coefs), *coefs, *pvecs)
        return self.loss(FBPINNJets.apply(theta, plan), all_params)

    def _fused_loss(self, ujss, all_params):
        loss = None
        for ic, weight, keys, data, fa, fb, static, dyn, done in self.fused:
            cols = []
            for k in keys:
                if k[0] == "J":
                    cols.append(ujss[ic][k[1]].reshape(-1))
                elif k[0] == "P":
                    fn, idx, perm, seg = data[k]
                    v = torch.as_tensor(fn(all_params)).reshape(-1).to(torch.float32)
                    cols.append(GatherRows.apply(v, idx, perm, seg))
                else:
                    cols.append(data[k])
            coef = self._coef(static, dyn, all_params)
            term = ResidualLoss.apply(coef, weight, fa, fb, done, *cols)
            loss = term if loss is None else loss + term
        return loss

    # -------------------------------------------------------------- per step

    def _ujs_generic(self, ic, q, all_params):
        off, n_c = self.cranges[ic]
        x0 = self.x[off:off + n_c]
        net = self.net
        C = self.problem.constraining_fn

        def u_fn(xb):
            d = xb - x0
            ut = q[0]
            for i in range(net.xd):
                if (net.d1mask >> i) & 1:
                    ut = ut + q[net.col1(i)] * d[:, i]
                if (net.d2mask >> i) & 1:
                    ut = ut + 0.5 * q[net.col2(i)] * d[:, i] * d[:, i]
            return C(all_params, xb, ut[:, None]), ()
        return get_ujs(x0, self.jmapss[ic], u_fn)

    def _req_codes(self, req):
        """fbpx_affine_jets request codes of one constraint's required_ujs."""
        xd = self.net.xd
        return tuple(0 if len(ixs) == 0 else (1 + ixs[0] if len(ixs) == 1 else 1 + xd + ixs[0])
                     for iu, ixs in req)

    def loss(self, ujet, all_params):
        """ujet: (J, n) jets of u~ at self.x.  Returns the scalar loss."""
        ujss = []
        if self.affine is not None:
            # Leibniz rule on the jets in one kernel per constraint
            A, B = self.affine
            reqs = [self._req_codes(r) for r in self.required_ujss]
            outs = AffineJets.apply(ujet, A, B, self.net.as_list(), list(self.cranges), reqs)
            i = 0
            for req in self.required_ujss:
                ujss.append(list(outs[i:i + len(req)]))
                i += len(req)
        else:
            for ic, (off, n_c) in enumerate(self.cranges):
                ujss.append(self._ujs_generic(ic, ujet[:, off:off + n_c], all_params))
        if self.fused is not None:
            return self._fused_loss(ujss, all_params)
        constraints_ = [list(self.cvalues[ic]) + ujss[ic] for ic in range(len(self.cranges))]
        return self.problem.loss_fn(all_params, constraints_)

    def u(self, ujet, all_params):
        """Constrained solution value u at self.x (value column only)."""
        q0 = ujet[0][:, None]
        return self.problem.constraining_fn(all_params, self.x, q0)


def _trainable_leaves(all_params):
    out = []

    def rec(d):
        if isinstance(d, dict):
            for v in d.values():
                rec(v)
        elif isinstance(d, (list, tuple)):
            for v in d:
                rec(v)
        elif isinstance(d, torch.Tensor) and d.requires_grad:
            out.append(d)
    rec(all_params["trainable"])
    return out


def _with_grad_trainable(all_params):
    """Shallow copy of all_params whose trainable leaves require grad."""
    def rec(d):
        if isinstance(d, dict):
            return {k: rec(v) for k, v in d.items()}
        if isinstance(d, (list, tuple)):
            return type(d)(rec(v) for v in d)
        if isinstance(d, torch.Tensor) and d.is_floating_point():
            return d.detach().double().requires_grad_(True)
        return d
    return {"static": all_params["static"], "trainable": rec(all_params["trainable"])}


# One launch for the POU reduction, residuals and transpose (fbpx_pou_loss)
# pays off where the points fit one workgroup; on larger grids it is slower
# than the separate launches, with the loads that do not depend on the
# reduction issued first too (steps/s: Burgers' first slab 5,635-5,684
# against 5,878-5,897, the batched inverse problem 11,867-11,999 against
# 12,269-12,274, Poisson 2,410-2,419 against 2,431-2,442).
POU_LOSS_MAX_POINTS = 256


class _JetsResidualLoss(torch.autograd.Function):
    """Loss of the fused residual description from the subdomain rows, as one
    autograd node (Epilogue.jets_loss): fbpx_forward -> fbpx_pou_reduce with
    the affine constraining function applied to every jet column ->
    fbpx_residual_loss per residual; backward: fbpx_residual_loss_backward
    writing d loss / d (constrained jets) straight into the epilogue's
    [J][n] buffer -> fbpx_pou_backward with the transposed affine map ->
    fbpx_backward -> fbpx_reduce_rows.  Same kernels and arithmetic as
    FBPINNJets + AffineJets + ResidualLoss, two launches fewer per step and
    no per-column gradient tensors.

    When theta's gradient is wanted and no residual coefficient is trained,
    the forward runs fbpx_residual_loss_grad instead: the same loss launch
    also writes d loss / d (constrained jets) at unit upstream gradient (and
    the per-point gradients of gathered columns), and the backward starts at
    fbpx_pou_backward, scaled by the incoming gradient there and in
    fbpx_segment_sum: one launch per residual fewer, the same values bit for
    bit at the step's unit seed.  When the residuals fit one fbpx_pou_loss
    launch (at most FBPX_POU_LOSS_RES residuals of at most
    FBPX_POU_LOSS_COLS columns) and the points one workgroup, that launch
    does the POU reduction, every residual and the POU transpose into the
    plan's gseg at unit upstream gradient, and the backward is fbpx_backward
    with the incoming gradient as its g_scale.  The shared buffers (the [J][n] jet gradients, the plan's
    gseg) carry an owner token; a backward that no longer owns them (another
    forward of this epilogue ran in between) takes the transpose path above."""

    @staticmethod
    def _columns(ep, C, i, vecs):
        """Residual i's factor columns (jet columns of C, data columns, and for
        a gathered column the trainable vector it reads, consumed from the
        iterator `vecs`), the gathered columns' positions and their indices."""
        ic, weight, keys, data, *_ = ep.fused[i]
        off, n_c = ep.cranges[ic]
        cols, gathered, idx = [], [], []
        for j, (c, k) in enumerate(zip(ep.combined[i], keys)):
            if c is not None:
                cols.append(C[c, off:off + n_c])
            elif k[0] == "P":
                cols.append(next(vecs))
                gathered.append(j)
                idx.append(data[k][1])
            else:
                cols.append(data[k])
        return cols, gathered, idx

    @staticmethod
    def forward(ctx, theta, plan, ep, nr, *inputs):
        coefs, pvecs = inputs[:nr], [v.contiguous() for v in inputs[nr:]]
        net, wl = plan.net, plan.wl
        ops.forward(net.as_list(), theta, plan.dom, plan.x, wl.tensor_list(), wl.size_list(), plan.pairbuf)
        A, B = ep.affine
        C = torch.empty((net.J, wl.n_points), dtype=torch.float32, device=theta.device)
        # the trainable vectors of the gathered columns (v[idx[k]] read in the
        # residual kernels), in key order
        pk = [data[k] for _, _, keys, data, *_ in ep.fused for k in keys if k[0] == "P"]
        # every residual adds its mean into one loss scalar
        total = torch.empty((), dtype=torch.float32, device=theta.device)
        # unit-seed gradients in the loss launch (see the class docstring)
        pregrad = ep._grad_wanted and ctx.needs_input_grad[0] and not any(ctx.needs_input_grad[4:4 + nr])
        fused_pou = pregrad and ep.pou_loss_fits and wl.n_points <= POU_LOSS_MAX_POINTS
        gp = []
        if fused_pou:
            rs = _JetsResidualLoss._pou_loss(ep, plan, C, coefs, pvecs, total, gp)
        elif ep.residual_set_fits:
            ops.pou_reduce(net.as_list(), wl.tensor_list(), wl.size_list(), plan.pairbuf, plan.wsum, C, A, B)
            rs = _JetsResidualLoss._residual_set(ep, C, coefs, pvecs, total, gp if pregrad else None)
        else:
            ops.pou_reduce(net.as_list(), wl.tensor_list(), wl.size_list(), plan.pairbuf, plan.wsum, C, A, B)
            rs = []
            it = iter(pvecs)
            for i, (ic, weight, keys, data, fa, fb, static, dyn, done) in enumerate(ep.fused):
                cols, gathered, idx = _JetsResidualLoss._columns(ep, C, i, it)
                r = torch.empty(ep.cranges[ic][1], dtype=torch.float32, device=theta.device)
                scratch = torch.empty(_lib.FBPX_RESIDUAL_BLOCKS, dtype=torch.float32, device=theta.device)
                if pregrad:
                    gidx, gcols = _JetsResidualLoss._gradient_columns(ep, i, gp)
                    ops.residual_loss_grad(cols, list(fa), list(fb), coefs[i], weight, r, scratch, total, done,
                                           gathered, idx, i > 0, gidx, gcols)
                else:
                    ops.residual_loss(cols, list(fa), list(fb), coefs[i], weight, r, scratch, total, done,
                                      gathered, idx, i > 0)
                rs.append(r)
        ctx.save_for_backward(theta, C, *coefs, *rs, *pvecs)
        ctx.plan, ctx.ep, ctx.nr, ctx.pk = plan, ep, nr, pk
        ctx.nvec = [v.numel() for v in pvecs]
        ctx.token = object() if pregrad else None
        ctx.fused_pou = fused_pou
        ctx.gp = gp
        ep._gC_owner = ctx.token
        return total

    @staticmethod
    def _residual_set(ep, C, coefs, pvecs, total, gp):
        """Every residual in one launch (fbpx_residual_losses), each as
        residual_loss_grad (gp given: the gradient columns at unit upstream
        gradient, gathered columns' buffers appended to gp) or residual_loss
        computes it.  Returns the residuals."""
        it = iter(pvecs)
        residuals, rs = [], []
        for i, (ic, weight, keys, data, fa, fb, static, dyn, done) in enumerate(ep.fused):
            cols, gathered, idx = _JetsResidualLoss._columns(ep, C, i, it)
            n_c = ep.cranges[ic][1]
            ix, gc = [None] * len(cols), [None] * len(cols)
            for j, t in zip(gathered, idx):
                ix[j] = t
            if gp is not None:
                for j, t in zip(*_JetsResidualLoss._gradient_columns(ep, i, gp)):
                    gc[j] = t
            rs.append(torch.empty(n_c, dtype=torch.float32, device=C.device))
            residuals.append(dict(n=n_c, weight=weight, coef=coefs[i], fa=fa, fb=fb, cols=cols, idx=ix, gcols=gc,
                                  r=rs[-1]))
        scratch = torch.empty(len(residuals) * _lib.FBPX_RESIDUAL_BLOCKS, dtype=torch.float32, device=C.device)
        ops.residual_losses(residuals, scratch, total, ep.fused[0][-1])
        return rs

    @staticmethod
    def _pou_loss(ep, plan, C, coefs, pvecs, total, gp):
        """fbpx_pou_loss over every residual: C, the residuals (returned), the
        gathered columns' per-point gradients (appended to gp) and plan.gseg
        at unit upstream gradient, and the loss into total."""
        net, wl = plan.net, plan.wl
        it = iter(pvecs)
        residuals, rs = [], []
        for i, (ic, weight, keys, data, fa, fb, static, dyn, done) in enumerate(ep.fused):
            off, n_c = ep.cranges[ic]
            cols, idx, gcols = [], [], []
            for c, k in zip(ep.combined[i], keys):
                if c is not None:
                    cols.append(int(c))
                    idx.append(None)
                    gcols.append(None)
                elif k[0] == "P":
                    cols.append(next(it))
                    idx.append(data[k][1])
                    gp.append(torch.empty(n_c, dtype=torch.float32, device=C.device))
                    gcols.append(gp[-1])
                else:
                    cols.append(data[k])
                    idx.append(None)
                    gcols.append(None)
            rs.append(torch.empty(n_c, dtype=torch.float32, device=C.device))
            residuals.append(dict(offset=off, n=n_c, weight=weight, coef=coefs[i], fa=fa, fb=fb, cols=cols, idx=idx,
                                  gcols=gcols, r=rs[-1]))
        nb = (wl.n_points + 255) // 256
        scratch = torch.empty(nb * len(residuals), dtype=torch.float32, device=C.device)
        A, B = ep.affine
        ops.pou_loss(net.as_list(), wl.tensor_list(), wl.size_list(), plan.pairbuf, plan.wsum, A, B, residuals, C,
                     plan.gseg, scratch, total, ep._pou_loss_done)
        return rs

    @staticmethod
    def _gradient_columns(ep, i, gp):
        """Residual i's gradient targets: the positions of its columns that
        take a gradient and their buffers (rows of ep.gC for jet columns, a new
        per-point buffer appended to gp for a gathered trainable vector)."""
        ic, _, keys, *_ = ep.fused[i]
        off, n_c = ep.cranges[ic]
        gidx, gcols = [], []
        for j, (c, k) in enumerate(zip(ep.combined[i], keys)):
            if c is not None:
                gidx.append(j)
                gcols.append(ep.gC[c, off:off + n_c])
            elif k[0] == "P":
                gidx.append(j)
                gcols.append(torch.empty(n_c, dtype=torch.float32, device=ep.gC.device))
                gp.append(gcols[-1])
        return gidx, gcols

    @staticmethod
    def backward(ctx, g):
        theta, C, *rest = ctx.saved_tensors
        plan, ep, nr, pk = ctx.plan, ctx.ep, ctx.nr, ctx.pk
        coefs, rs, pvecs = rest[:nr], rest[nr:2 * nr], rest[2 * nr:]
        net, wl = plan.net, plan.wl
        g = g.contiguous()
        # the forward's unit-seed gradients are still in ep.gC (or, from
        # fbpx_pou_loss, in plan.gseg): scale by g where they are read; else
        # the transposes write them here
        pregrad = ctx.token is not None and ep._gC_owner is ctx.token
        ep._gC_owner = None
        scale = g if pregrad else None
        gp = ctx.gp if pregrad else []   # gp: per gathered column, its per-point gradient buffer
        gcoefs = [None] * nr
        it = iter(pvecs)
        for i, (ic, weight, keys, data, fa, fb, static, dyn, done) in enumerate(ep.fused if not pregrad else ()):
            cols, gathered, idx = _JetsResidualLoss._columns(ep, C, i, it)
            gidx, gcols = _JetsResidualLoss._gradient_columns(ep, i, gp)
            want = bool(ctx.needs_input_grad[4 + i])
            nt = len(fa)
            gcoef = torch.empty(nt, dtype=torch.float32, device=g.device)
            scratch = torch.empty(_lib.FBPX_RESIDUAL_BLOCKS * nt, dtype=torch.float32, device=g.device)
            ops.residual_loss_backward(cols, list(fa), list(fb), coefs[i], rs[i], weight, g, gidx, gcols, scratch,
                                       gcoef, want, done, gathered, idx)
            gcoefs[i] = gcoef if want else None
        # transposes of the gathers: per trainable vector, its points summed in
        # index order; a single one shares the POU transpose's launch
        gvecs = [torch.empty(nv, dtype=torch.float32, device=g.device) for nv in ctx.nvec]
        with_pou = not (pregrad and ctx.fused_pou) and len(gvecs) == 1
        for j, (d, gcol, out) in enumerate(zip(pk, gp, gvecs)):
            if not with_pou:
                ops.segment_sum(d[3], d[2], gcol, out, scale)
        if pregrad and ctx.fused_pou:
            _JetsResidualLoss._backward(plan, theta, g)
        else:
            if with_pou:
                d = pk[0]
                ops.pou_backward_segment_sum(net.as_list(), wl.tensor_list(), wl.size_list(), plan.wsum, ep.gC,
                                             plan.gseg, ep.affine[1], scale, d[3], d[2], gp[0], gvecs[0])
            else:
                ops.pou_backward(net.as_list(), wl.tensor_list(), wl.size_list(), plan.wsum, ep.gC, plan.gseg,
                                 ep.affine[1], scale)
            _JetsResidualLoss._backward(plan, theta, None)
        return (ops.theta_gradient(plan, theta), None, None, None, *gcoefs, *gvecs)

    @staticmethod
    def _backward(plan, theta, g_scale):
        """fbpx_backward into plan.partial, or, in the engine's step when it
        allows it (plan.adam_fold), with the optimizer step folded in."""
        net, wl = plan.net, plan.wl
        if plan.defer_rows and plan.adam_fold is not None:
            th, mu, nu, count, count_done, lr, b1, b2, eps = plan.adam_fold
            ops.backward_adam(net.as_list(), th, plan.dom, plan.x, wl.tensor_list(), wl.size_list(),
                              plan.chunk_train, plan.gseg, g_scale, mu, nu, count, count_done, lr, b1, b2, eps)
            plan.adam_folded = True
        else:
            ops.backward(net.as_list(), theta, plan.dom, plan.x, wl.tensor_list(), wl.size_list(), plan.chunk_train,
                         plan.gseg, plan.partial, g_scale)
