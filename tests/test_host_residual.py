"""CPU: the residual descriptions of the built-in problems and the set-up
check that decides between the fused residual loss and loss_fn.

Each description agrees with its loss_fn (float64, random derivatives, value
and gradients); descriptions with a planted mistake (a flipped coefficient, a
wrong derivative, a dropped term, a wrong weight, a wrong trainable
coefficient) are rejected.
"""

import dataclasses

import pytest
import torch

import setups
from fbpinns_amd import epilogue
from fbpinns_amd import problems as P
from fbpinns_amd.engine import pick_schema
from fbpinns_amd.ops import NetSpec
from fbpinns_amd.residuals import J, Residual, V, check_spec
from fbpinns_amd.residuals import P as PF

PHYS1 = ((0, ()), (0, (0,)), (0, (0, 0)))
CASES = {
    "ho_soft": (P.HarmonicOscillator1D, dict(d=2, w0=20), 1, [PHYS1, ((0, ()), (0, (0,)))], [0, 2], [40, 1]),
    "ho_hard": (P.HarmonicOscillator1DHardBC, dict(d=2, w0=20), 1, [PHYS1], [0], [40]),
    "ho_inverse": (P.HarmonicOscillator1DInverse, dict(d=2, w0=20), 1, [PHYS1, ((0, ()),)], [0, 1], [40, 13]),
    "burgers": (P.BurgersEquation2D, dict(), 2, [((0, ()), (0, (0,)), (0, (1,)), (0, (0, 0)))], [0], [50]),
    "wave": (P.WaveEquationConstantVelocity3D, dict(c0=1.3, source=P.wave_sources()), 3,
             [((0, (0, 0)), (0, (1, 1)), (0, (2, 2)))], [0], [50]),
    "wave_gauss": (setups.VariableVelocityWave3D, dict(c0=1.0, source=P.wave_sources()), 3,
                   [((0, (0, 0)), (0, (1, 1)), (0, (2, 2)))], [0], [50]),
    "poisson": (P.PoissonMultiscale2D, dict(), 2, [((0, (0, 0)), (0, (1, 1)))], [0], [50]),
    "batched_inverse": (P.BatchedHarmonicOscillatorInverse,
                        dict(w0s=[20.0, 35.0, 50.0, 80.0], mus_true=[1.0, 2.0, 4.0, 6.0], n_obs=4), 1,
                        [PHYS1, ((0, ()),)], [["i"], ["f", "i"]], [40, 16]),
}


def _value(kind, n, g, n_inst=4):
    if kind == "i":
        return torch.randint(0, n_inst, (n,), generator=g)
    return torch.randn((n, 1), generator=g, dtype=torch.float64)


def _setup(name):
    prob, kw, xd, reqs, nvals, sizes = CASES[name]
    s, t = prob.init_params(**kw)
    ap = {"static": {"problem": s},
          "trainable": {"problem": {k: 0.7 + 0.1 * torch.arange(v.numel(), dtype=v.dtype).reshape(v.shape)
                                    for k, v in t.items()}}}
    g = torch.Generator().manual_seed(0)
    cranges, off = [], 0
    for n in sizes:
        cranges.append((off, n))
        off += n
    x = torch.rand((off, xd), generator=g, dtype=torch.float64) * 0.8 + 0.1
    cvalues = [[x[o:o + n]] + [_value(k, n, g) for k in (["f"] * nv if isinstance(nv, int) else nv)]
               for (o, n), nv in zip(cranges, nvals)]
    d1, d2 = pick_schema(xd, *epilogue.schema_masks(reqs, xd))
    ep = epilogue.Epilogue(prob, ap, x, cranges, cvalues, reqs, NetSpec(xd, 32, 2, d1, d2))
    return prob, ap, ep, [len(c) - 1 for c in cvalues], reqs


@pytest.mark.parametrize("name", list(CASES))
def test_description_agrees_with_loss_fn(name):
    prob, ap, ep, nvals, reqs = _setup(name)
    spec = prob.residuals(ap)
    check_spec(spec, len(reqs), [len(r) for r in reqs], nvals)
    assert ep._fused_agrees(spec, ap, nvals)


def _mutations(spec, n_req):
    """Descriptions that differ from spec by one plausible mistake."""
    out = []
    for ir, res in enumerate(spec):
        for it, (c, a, b) in enumerate(res.terms):
            terms = list(res.terms)
            if callable(c):
                terms[it] = (lambda p, c=c: 1.01 * c(p), a, b)
            else:
                terms[it] = (-c if c != 0 else 1.0, a, b)
            out.append(("coef", ir, it, terms))
            if isinstance(a, PF):
                terms = list(res.terms)
                terms[it] = (c, PF(lambda p, fn=a.fn: torch.roll(fn(p), 1), a.j), b)
                out.append(("gather", ir, it, terms))
            if isinstance(a, J) and n_req[res.constraint] > 1:
                terms = list(res.terms)
                terms[it] = (c, J((a.i + 1) % n_req[res.constraint]), b)
                out.append(("factor", ir, it, terms))
            if len(res.terms) > 1:
                terms = list(res.terms)
                del terms[it]
                out.append(("drop", ir, it, terms))
        out.append(("weight", ir, -1, None))
    specs = []
    for kind, ir, it, terms in out:
        s2 = list(spec)
        if kind == "weight":
            s2[ir] = dataclasses.replace(spec[ir], weight=spec[ir].weight * 2)
        else:
            s2[ir] = dataclasses.replace(spec[ir], terms=terms)
        specs.append((f"{kind} r{ir} t{it}", s2))
    return specs


@pytest.mark.parametrize("name", list(CASES))
def test_planted_mistakes_are_rejected(name):
    prob, ap, ep, nvals, reqs = _setup(name)
    spec = prob.residuals(ap)
    muts = _mutations(spec, [len(r) for r in reqs])
    assert muts
    for label, s2 in muts:
        assert not ep._fused_agrees(s2, ap, nvals), label


def test_check_spec_rejects_bad_structure():
    good = [Residual(0, 1.0, [(1.0, J(0), None)])]
    check_spec(good, 1, [1], [0])
    for bad in ([Residual(1, 1.0, [(1.0, J(0), None)])],       # constraint out of range
                [Residual(0, 1.0, [(1.0, J(1), None)])],       # derivative out of range
                [Residual(0, 1.0, [(1.0, V(0), None)])],       # no constraining value
                [Residual(0, 1.0, [])],                        # no terms
                [Residual(0, 1.0, [(1.0, J(0), None)] * 9)],   # too many terms
                [Residual(0, 1.0, [("a", J(0), None)])],       # coefficient type
                []):
        with pytest.raises(ValueError):
            check_spec(bad, 1, [1], [0])


def test_problem_without_description_runs_loss_fn():
    assert P.Problem.residuals({}) is None


@pytest.mark.parametrize("name", list(CASES))
def test_fused_plans_structure(name):
    """Per residual: every term factor names a column slot, derivative slots
    index the constraint's requests, per-point data columns are float32 (n,),
    gathers carry a permutation that sorts their index and segment offsets
    that count it."""
    prob, ap, ep, nvals, reqs = _setup(name)
    spec = prob.residuals(ap)
    plans = ep._fused_plans(spec, ap)
    assert len(plans) == len(spec)
    for (ic, weight, keys, data, fa, fb, static, dyn, done), res in zip(plans, spec):
        assert done.dtype == torch.int32 and int(done.item()) == 0
        n_c = ep.cranges[ic][1]
        assert ic == res.constraint and weight == res.weight
        assert len(fa) == len(fb) == len(res.terms) == static.numel()
        assert len(set(keys)) == len(keys)
        for f in fa + fb:
            assert -1 <= f < len(keys)
        for k in keys:
            if k[0] == "J":
                assert 0 <= k[1] < len(reqs[ic])
            elif k[0] == "P":
                fn, idx, perm, seg = data[k]
                assert idx.dtype == perm.dtype == seg.dtype == torch.int32
                srt = idx[perm.long()]
                assert (srt[1:] >= srt[:-1]).all()
                assert int(seg[-1]) == n_c and torch.equal(seg[1:] - seg[:-1], torch.bincount(
                    idx.long(), minlength=seg.numel() - 1).to(torch.int32))
            else:
                assert data[k].dtype == torch.float32 and data[k].shape == (n_c,)
        for t, (c, a, b) in enumerate(res.terms):
            assert (t in [d[0] for d in dyn]) == callable(c)
            if not callable(c):
                assert float(static[t]) == pytest.approx(float(c), rel=1e-7)
