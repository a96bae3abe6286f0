"""Pointwise-residual form of a problem's loss, for the fused loss kernel.

Every built-in loss_fn of the reference (problems.py:123-124, 204, 265, 326,
405) is a weighted sum of means of squared pointwise residuals, each residual
a short sum of products of the constraint's derivatives and per-point data:

    loss = sum over residuals of  weight * mean_k( sum_t coef_t * a_t[k] * b_t[k] )^2.

A problem may describe its loss that way with a static method

    residuals(all_params) -> [Residual(constraint, weight, terms), ...] or None

where each term is (coef, a, b):
  coef  a float, or a callable all_params -> 0-d tensor (trainable
        coefficients such as a learnt damping);
  a, b  None (factor 1), J(i) (the i-th derivative the constraint requests,
        in its required_ujs order), V(j) (the j-th constraining value of the
        constraint, a per-point (n, 1) tensor), F(fn) (per-point data
        fn(all_params, x_c, *values_c) -> (n, 1) or (1, 1), from static
        parameters only: it is evaluated once per work list) or P(fn, j) (a
        trainable vector fn(all_params) gathered per point by the j-th
        constraining value, an integer index: per-instance coefficients).

The engine then evaluates each residual with one fused kernel pair
(fbpx_residual_loss, and its transpose in the backward) instead of the chain
of elementwise ops and the mean of loss_fn.  The description is checked
against loss_fn at set-up (float64, random derivatives: value and gradients
with respect to the derivatives and the trainable parameters); a problem
whose description disagrees runs loss_fn itself.
"""

from dataclasses import dataclass
from typing import Callable, List, Optional, Tuple, Union

import torch

from fbpinns_amd import _lib


@dataclass(frozen=True)
class J:
    """The i-th requested derivative of the residual's constraint."""
    i: int


@dataclass(frozen=True)
class V:
    """The j-th constraining value of the residual's constraint."""
    j: int


@dataclass(frozen=True)
class F:
    """Per-point data fn(all_params, x_c) from static parameters."""
    fn: Callable


@dataclass(frozen=True)
class P:
    """Trainable vector fn(all_params) gathered by the j-th constraining value."""
    fn: Callable
    j: int


Factor = Optional[Union[J, V, F, P]]
Coef = Union[float, Callable]


@dataclass
class Residual:
    constraint: int
    weight: float
    terms: List[Tuple[Coef, Factor, Factor]]


def check_spec(spec, n_constraints, n_req, n_values):
    """Structural check of a residual list; raises ValueError."""
    if not isinstance(spec, (list, tuple)) or not spec:
        raise ValueError("residuals() must return a non-empty list of Residual")
    for res in spec:
        if not isinstance(res, Residual):
            raise ValueError(f"not a Residual: {res!r}")
        ic = res.constraint
        if not 0 <= ic < n_constraints:
            raise ValueError(f"residual over constraint {ic} of {n_constraints}")
        if not 1 <= len(res.terms) <= _lib.FBPX_MAX_TERMS:
            raise ValueError(f"residual with {len(res.terms)} terms (1 to {_lib.FBPX_MAX_TERMS})")
        for coef, a, b in res.terms:
            if not (isinstance(coef, (int, float)) or callable(coef)):
                raise ValueError(f"coefficient {coef!r} is neither a number nor a callable")
            for f in (a, b):
                if f is None or isinstance(f, F):
                    continue
                if isinstance(f, J) and 0 <= f.i < n_req[ic]:
                    continue
                if isinstance(f, (V, P)) and 0 <= f.j < n_values[ic]:
                    continue
                raise ValueError(f"factor {f!r} outside constraint {ic} "
                                 f"({n_req[ic]} derivatives, {n_values[ic]} values)")


def spec_loss(spec, all_params, constraints, n_values):
    """The loss a residual list describes, in torch (set-up check).

    constraints: per constraint [x_c, *values_c, *derivatives_c], as passed to
    loss_fn."""
    total = None
    for res in spec:
        c = constraints[res.constraint]
        x = c[0]
        nv = n_values[res.constraint]

        def col(f):
            if f is None:
                return 1.0
            if isinstance(f, J):
                return c[1 + nv + f.i]
            if isinstance(f, V):
                return c[1 + f.j]
            if isinstance(f, P):
                return f.fn(all_params).reshape(-1)[c[1 + f.j].reshape(-1).long()].reshape(-1, 1)
            return f.fn(all_params, x, *c[1:1 + nv])
        r = 0.0
        for coef, a, b in res.terms:
            cf = coef(all_params) if callable(coef) else coef
            r = r + cf * col(a) * col(b)
        term = res.weight * torch.mean(r ** 2)
        total = term if total is None else total + term
    return total
