"""The benchmark configurations, built through the public Constants API.

ho       damped oscillator, hard BC, 15 subdomains of width 0.15, [1,32,1],
         200 grid points, d=2, w0=80 (README.md:123-148 with the hard-BC
         problem, problems.py:149-206)
poisson  multiscale Poisson on [0,1]^2, 16x16 subdomains, overlap 1.9,
         [2,32,32,1], 262,144 uniform points (seed 1)
wave     (2+1)D wave, c=1, three Gaussian sources (sigma 0.3, 0.1, 0.03),
         12x12x6 subdomains, overlap 1.9, [3,64,64,1], 1,048,576 uniform
         points (seed 2)
burgers  (1+1)D Burgers, 30x6 subdomains, overlap 2.9, [2,32,32,1], 400x400
         grid, time slabs of 2,000 steps
"""

import numpy as np

from fbpinns_amd import decompositions, domains, networks, problems, schedulers
from fbpinns_amd.constants import Constants, get_subdomain_ws


def _rect(lo, hi, nsub, overlap=None, widths=None):
    xs = [np.linspace(a, b, n) for a, b, n in zip(lo, hi, nsub)]
    ws = widths if widths is not None else get_subdomain_ws(xs, overlap)
    return dict(subdomain_xs=xs, subdomain_ws=ws, unnorm=(0.0, 1.0))


def ho(**kw):
    c = Constants(
        run="ho_hard_w80",
        domain=domains.RectangularDomainND,
        domain_init_kwargs=dict(xmin=np.array([0.0]), xmax=np.array([1.0])),
        problem=problems.HarmonicOscillator1DHardBC,
        problem_init_kwargs=dict(d=2, w0=80, sd=0.1),
        decomposition=decompositions.RectangularDecompositionND,
        decomposition_init_kwargs=_rect((0.0,), (1.0,), (15,), widths=[0.15 * np.ones(15)]),
        network=networks.FCN,
        network_init_kwargs=dict(layer_sizes=[1, 32, 1], init="glorot_normal"),
        ns=((200,),), n_test=(500,), n_steps=1000, sampler="grid",
        optimiser_kwargs=dict(learning_rate=1e-3), seed=0,
        summary_freq=250, test_freq=250, model_save_freq=10 ** 9)
    for k, v in kw.items():
        c[k] = v
    return c


def poisson(**kw):
    c = Constants(
        run="poisson_multiscale",
        domain=domains.RectangularDomainND,
        domain_init_kwargs=dict(xmin=np.array([0.0, 0.0]), xmax=np.array([1.0, 1.0])),
        problem=problems.PoissonMultiscale2D,
        problem_init_kwargs=dict(ks=(1, 2, 3, 4), sd=0.1),
        decomposition=decompositions.RectangularDecompositionND,
        decomposition_init_kwargs=_rect((0.0, 0.0), (1.0, 1.0), (16, 16), 1.9),
        network=networks.FCN,
        network_init_kwargs=dict(layer_sizes=[2, 32, 32, 1]),
        ns=((512, 512),), n_test=(200, 200), n_steps=5000, sampler="uniform",
        optimiser_kwargs=dict(learning_rate=1e-3), seed=1,
        summary_freq=500, test_freq=1000, model_save_freq=10 ** 9)
    for k, v in kw.items():
        c[k] = v
    return c


def wave(**kw):
    c = Constants(
        run="wave_3gauss",
        domain=domains.RectangularDomainND,
        domain_init_kwargs=dict(xmin=np.array([-1.0, -1.0, 0.0]), xmax=np.array([1.0, 1.0, 1.0])),
        problem=problems.WaveEquationConstantVelocity3D,
        problem_init_kwargs=dict(c0=1, source=problems.wave_sources(seed=2)),
        decomposition=decompositions.RectangularDecompositionND,
        decomposition_init_kwargs=_rect((-1.0, -1.0, 0.0), (1.0, 1.0, 1.0), (12, 12, 6), 1.9),
        network=networks.FCN,
        network_init_kwargs=dict(layer_sizes=[3, 64, 64, 1]),
        ns=((128, 128, 64),), n_test=(100, 100, 10), n_steps=10000, sampler="uniform",
        optimiser_kwargs=dict(learning_rate=1e-3), seed=2,
        summary_freq=1000, test_freq=5000, model_save_freq=10 ** 9)
    for k, v in kw.items():
        c[k] = v
    return c


def burgers(**kw):
    c = Constants(
        run="burgers_slabs",
        domain=domains.RectangularDomainND,
        domain_init_kwargs=dict(xmin=np.array([-1.0, 0.0]), xmax=np.array([1.0, 1.0])),
        problem=problems.BurgersEquation2D,
        problem_init_kwargs=dict(),
        decomposition=decompositions.RectangularDecompositionND,
        decomposition_init_kwargs=_rect((-1.0, 0.0), (1.0, 1.0), (30, 6), 2.9),
        network=networks.FCN,
        network_init_kwargs=dict(layer_sizes=[2, 32, 32, 1]),
        ns=((400, 400),), n_test=(200, 100), n_steps=12000, sampler="grid",
        scheduler=schedulers.SlabSchedulerRectangularND,
        scheduler_kwargs=dict(axis=1, steps_per_slab=2000),
        optimiser_kwargs=dict(learning_rate=1e-3), seed=0,
        summary_freq=1000, test_freq=2000, model_save_freq=10 ** 9)
    for k, v in kw.items():
        c[k] = v
    return c


def inverse_batch(n_inst=256, seed=3, **kw):
    """n_inst independent inverse oscillators in one launch: w0 ~ U(20, 100),
    true mu ~ U(1, 6), subdomain count per instance ~ U{5..40} (seed 3),
    widths 2.1 x centre spacing, [1,32,1], 200 grid collocation points and 64
    noisy (sigma 0.01) observations per instance, learnable mu_i from 0."""
    from fbpinns_amd.util import rng
    k1, k2, k3 = rng.split(rng.PRNGKey(seed), 3)
    w0s = rng.uniform(k1, (n_inst,), 20.0, 100.0)
    mus = rng.uniform(k2, (n_inst,), 1.0, 6.0)
    ms = np.floor(rng.uniform(k3, (n_inst,), 5.0, 41.0)).astype(int)
    spacing = 2.0
    xss = [[np.linspace(0.0, 1.0, m) + spacing * i] for i, m in enumerate(ms)]
    wss = [get_subdomain_ws(xs, 2.1) for xs in xss]
    c = Constants(
        run="inverse_batch",
        domain=domains.InstancesDomain1D,
        domain_init_kwargs=dict(n_inst=n_inst, spacing=spacing),
        problem=problems.BatchedHarmonicOscillatorInverse,
        problem_init_kwargs=dict(w0s=w0s, mus_true=mus, n_obs=64, noise=0.01, seed=seed, spacing=spacing),
        decomposition=decompositions.ConcatenatedRectangularDecompositionND,
        decomposition_init_kwargs=dict(subdomain_xss=xss, subdomain_wss=wss, unnorm=(0.0, 1.0)),
        network=networks.FCN,
        network_init_kwargs=dict(layer_sizes=[1, 32, 1]),
        ns=((200,),), n_test=(200,), n_steps=20000, sampler="grid",
        optimiser_kwargs=dict(learning_rate=1e-3), seed=seed,
        summary_freq=5000, test_freq=5000, model_save_freq=10 ** 9)
    for k, v in kw.items():
        c[k] = v
    return c


CONFIGS = {"ho": ho, "poisson": poisson, "wave": wave, "burgers": burgers, "inverse": inverse_batch}

# forward FLOPs per (point, subdomain) pair for a [xd, h x nh, 1] network
# carrying J jet columns: layer 1 value column only (its tangents are
# analytic), every later contraction for all J columns.
def fwd_flops_per_pair(xd, h, nh, J):
    f = 2 * xd * h
    if nh == 2:
        f += 2 * h * h * J
    f += 2 * h * J
    return f


# matrix-core FLOPs the kernels execute per (point, subdomain) pair: the
# layer-2 contractions run as passes of v_mfma_f32_16x16x4f32 over 16 pairs,
# (h/16) x (h/4) instructions of 2 * 16 * 16 * 4 FLOPs per pass.  Forward: Z_0,
# the W2^T vector when any first derivative is carried, one Z_i per axis with
# a second derivative.  Backward: Z_0; with first derivatives the s-vector
# pass, its W2^T pass and its dW2 column; per second-derivative axis Z_i, its
# W2^T pass and dW2 column; W2^T G_0 and the G_0 column.  One hidden layer:
# no matrix-core work.
def mfma_flops_per_pair(h, nh, n1, n2):
    if nh != 2:
        return 0, 0
    per_pass = (h // 16) * (h // 4) * 2 * 16 * 16 * 4 / 16
    fwd = (1 + (n1 > 0) + n2) * per_pass
    bwd = (3 + 3 * (n1 > 0) + 3 * n2) * per_pass
    return fwd, bwd
