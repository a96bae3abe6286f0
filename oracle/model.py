"""Oracle: FBPINN model, derivative jets and loss (torch on CPU, float64/float32).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Restates, with the same structure as the reference (gather -> vmap over
(point, subdomain) pairs -> two segment sums -> constraining_fn, and one
nested forward-mode jvp per leaf of the derivative tree):
  * get_jmaps              trainers.py:63-107
  * FBPINN_model_inner     trainers.py:113-118 with
      RectangularDecompositionND.norm_fn/unnorm_fn/window_fn decompositions.py:183-199,
      networks.norm/unnorm networks.py:197-201, windows.cosine windows.py:25-35,
      FCN.network_fn networks.py:60-68
  * FBPINN_model           trainers.py:126-177
  * FBPINN_forward/_get_ujs/jacfwd trainers.py:197-247
  * FBPINN_loss            trainers.py:249-267
  * value_and_grad of the loss (trainers.py:292) via torch.autograd

Pinned by (tests/test_oracle_model.py): central finite differences in float64
of the model value (first and second derivatives) and of the loss w.r.t.
parameters; the POU invariant (every subdomain emitting the same constant c
gives u == c and zero derivatives); the m=1, flag=0 reduction to a plain PINN
evaluated with an explicit hand-written tanh-MLP derivative recursion; the
closed-form window value/derivatives at the box centre and edges.
"""

import math

import numpy as np
import torch
from torch.func import jvp, vmap


# ---------------------------------------------------------------- derivative tree

def get_jmaps(required_ujs):
    """trainers.py:63-107: nodes, leaves and jac_is for the nested-jvp tree."""
    tree = {}
    for iu, ixs in required_ujs:
        t = tree
        for ix in ixs:
            if ix not in t:
                t[ix] = {}
            t = t[ix]

    def get_nodes(t, n=(), ks=()):
        ni = len(n)
        for k in t:
            ks_ = ks + (k,)
            if t[k]:
                n += (((ni, k), ks_, 0),)
                n = get_nodes(t[k], n, ks_)
            else:
                n += (((ni, k), ks_, 1),)
        return n

    nodes = get_nodes(tree)
    leaves = tuple((i + 1, node[1]) for i, node in enumerate(nodes) if node[2])
    if not leaves:
        leaves = ((0, ()),)
    jac_is = ()
    for iu, ixs in required_ujs:
        io = len(ixs)
        il = [leaf[1][:io] for leaf in leaves].index(ixs)
        jac_is += ((il, io, iu),)
    return nodes, leaves, jac_is


def _jacfwd(f, v):
    """trainers.py:241-247."""
    def jacfun(x):
        y, j, aux = jvp(f, (x,), (v,), has_aux=True)
        return j, aux + (y,)
    return jacfun


def get_ujs(x_batch, jmaps, u):
    """trainers.py:213-239.  `u(x_batch) -> ((n, ud), ())`."""
    nodes, leaves, jac_is = jmaps
    n, xd = x_batch.shape
    vs = torch.eye(xd, dtype=x_batch.dtype).expand(n, xd, xd)
    fs = [u]
    for (ni, ix), _, _ in nodes:
        fs.append(_jacfwd(fs[ni], vs[:, ix].contiguous()))
    jacs = []
    for ie, _ in leaves:
        fin, jac = fs[ie](x_batch)
        jacs.append(jac + (fin,))
    return [jacs[il][io][:, iu:iu + 1] for il, io, iu in jac_is]


# ---------------------------------------------------------------- model

def _pair_fn(Ws, bs, xmin, xmax, flag, unnorm, x):
    """FBPINN_model_inner for ONE (point, subdomain) pair (trainers.py:113-118)."""
    mu, sd = (xmax + xmin) / 2, (xmax - xmin) / 2           # decompositions.py:187
    h = (x - mu) / sd                                        # networks.py:198
    for W, b in zip(Ws[:-1], bs[:-1]):                       # networks.py:63-65
        h = torch.tanh(W @ h + b)
    u_raw = Ws[-1] @ h + bs[-1]                              # networks.py:66-67
    u = u_raw * unnorm[1] + unnorm[0]                        # networks.py:201
    wmu, wsd = (xmin + xmax) / 2, (xmax - xmin) / 2          # windows.py:28
    ws = ((1 + torch.cos(math.pi * (x - wmu) / wsd)) / 2) ** 2
    # heaviside(., 1): 1 where the argument is >= 0, zero derivative
    ws = (x >= xmin).to(x.dtype) * (xmax >= x).to(x.dtype) * ws
    w = torch.prod(ws, dim=0, keepdim=True)                  # windows.py:33
    w = flag * w + (1 - flag)                                # decompositions.py:199
    return u * w, w, u_raw


def fbpinn_model(layers, dom, x_batch, takes, constraining_fn):
    """trainers.py:126-177.

    layers : list of (W (m, out, in), b (m, out)) indexed by m_take
    dom    : dict xmins, xmaxs (m, xd), flags (m, 1), unnorms (m, 2)
    takes  : (m_take, n_take, p_take, np_take, npou)
    constraining_fn(x_batch, u) -> u
    Returns u (n, ud), wp, us, ws, us_raw.
    """
    m_take, n_take, p_take, np_take, npou = takes
    m_take = torch.as_tensor(m_take, dtype=torch.long)
    n_take = torch.as_tensor(n_take, dtype=torch.long)
    p_take = torch.as_tensor(p_take, dtype=torch.long)
    np_take = torch.as_tensor(np_take, dtype=torch.long)
    x_take = x_batch[n_take]
    Ws = [W[m_take] for W, _ in layers]
    bs = [b[m_take] for _, b in layers]
    args = (Ws, bs, dom["xmins"][m_take], dom["xmaxs"][m_take],
            dom["flags"][m_take], dom["unnorms"][m_take], x_take)
    us, ws, us_raw = vmap(_pair_fn)(*args)
    u = torch.cat([us, ws], dim=1)
    seg = torch.zeros((len(np_take), u.shape[1]), dtype=u.dtype).index_add(0, p_take, u)
    wp = seg[:, -1:]
    u = seg[:, :-1] / wp
    u = torch.zeros((x_batch.shape[0], u.shape[1]), dtype=u.dtype).index_add(0, np_take, u)
    u = u / npou
    u = constraining_fn(x_batch, u)
    return u, wp, us, ws, us_raw


def fbpinn_forward(layers, dom, x_batch, takes, constraining_fn, jmaps):
    """trainers.py:197-203."""
    def u(xb):
        return fbpinn_model(layers, dom, xb, takes, constraining_fn)[0], ()
    return get_ujs(x_batch, jmaps, u)


def fbpinn_loss(layers, dom, constraints, takess, jmapss, constraining_fn, loss_fn):
    """trainers.py:249-267.  constraints: list of [x_batch, *values];
    loss_fn(constraints_) with constraints_[i] = constraint + ujs."""
    cs = []
    for takes, jmaps, constraint in zip(takess, jmapss, constraints):
        ujs = fbpinn_forward(layers, dom, constraint[0], takes, constraining_fn, jmaps)
        cs.append(list(constraint) + ujs)
    return loss_fn(cs)


def identity_constraint(x_batch, u):
    return u


# ---------------------------------------------------------------- helpers

def to_torch_layers(layers, dtype):
    return [(torch.as_tensor(np.asarray(W), dtype=dtype), torch.as_tensor(np.asarray(b), dtype=dtype))
            for W, b in layers]


def to_torch_dom(xmins, xmaxs, flags, unnorms, dtype):
    c = lambda a: torch.as_tensor(np.asarray(a), dtype=dtype)
    return dict(xmins=c(xmins), xmaxs=c(xmaxs), flags=c(flags), unnorms=c(unnorms))


def value_and_grad(layers, dom, constraints, takess, jmapss, constraining_fn, loss_fn,
                   active_rows=None, extra=()):
    """Loss and gradients w.r.t. every layer tensor (rows = subdomains) and
    any extra trainable tensors (trainers.py:292 value_and_grad)."""
    layers = [(W.detach().clone().requires_grad_(True), b.detach().clone().requires_grad_(True))
              for W, b in layers]
    extra = [e.detach().clone().requires_grad_(True) for e in extra]
    loss = fbpinn_loss(layers, dom, constraints, takess, jmapss, constraining_fn,
                       (lambda cs: loss_fn(cs, *extra)) if extra else loss_fn)
    flat = [t for W, b in layers for t in (W, b)] + extra
    grads = torch.autograd.grad(loss, flat, allow_unused=True)
    grads = [torch.zeros_like(t) if g is None else g for g, t in zip(grads, flat)]
    nl = len(layers)
    glayers = [(grads[2 * i], grads[2 * i + 1]) for i in range(nl)]
    return loss.detach(), glayers, grads[2 * nl:]
