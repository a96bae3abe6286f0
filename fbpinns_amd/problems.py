"""PDE problems (reference: problems.py), written with torch ops.

Each problem keeps the reference's static-method API: init_params,
sample_constraints, constraining_fn, loss_fn, exact_solution.  Means inside
loss_fn use fbpinns_amd.ops.mean (a deterministic, graph-capture-safe
reduction kernel) in place of jnp.mean.
"""

import math

import numpy as np
import torch

from fbpinns_amd.ops import mean
from fbpinns_amd.residuals import F, J, P, Residual, V
from fbpinns_amd.traditional_solutions.burgers import burgers_viscous_time_exact1
from fbpinns_amd.traditional_solutions.seismic_cpml import seismicCPML2D
from fbpinns_amd.util import rng
from fbpinns_amd.util.logger import logger


class Problem:
    """Base problem class (problems.py:19-62)."""

    @staticmethod
    def init_params(*args):
        raise NotImplementedError

    @staticmethod
    def sample_constraints(all_params, domain, key, sampler, batch_shapes):
        """Returns [[x_batch, *constraining_values, required_ujs], ...]."""
        raise NotImplementedError

    @staticmethod
    def constraining_fn(all_params, x_batch, u):
        return u

    @staticmethod
    def loss_fn(all_params, constraints):
        raise NotImplementedError

    @staticmethod
    def residuals(all_params):
        """Optional description of loss_fn as weighted means of squared
        pointwise residuals (fbpinns_amd.residuals), which the engine then
        evaluates with the fused residual-loss kernels; None: loss_fn runs."""
        return None

    @staticmethod
    def exact_solution(all_params, x_batch, batch_shape=None):
        raise NotImplementedError


def _t(a, like):
    return torch.as_tensor(a, dtype=like.dtype, device=like.device)


class HarmonicOscillator1D(Problem):
    """m u'' + mu u' + k u = 0, u(0)=1, u'(0)=0 with soft BCs (problems.py:68-146)."""

    @staticmethod
    def init_params(d=2, w0=20):
        mu, k = 2 * d, w0 ** 2
        return {"dims": (1, 1), "d": d, "w0": w0, "mu": mu, "k": k}, {}

    @staticmethod
    def sample_constraints(all_params, domain, key, sampler, batch_shapes):
        x_phys = domain.sample_interior(all_params, key, sampler, batch_shapes[0])
        req_phys = ((0, ()), (0, (0,)), (0, (0, 0)))
        x_bd = torch.tensor([[0.0]])
        u_bd = torch.tensor([[1.0]])
        ut_bd = torch.tensor([[0.0]])
        req_bd = ((0, ()), (0, (0,)))
        return [[x_phys, req_phys], [x_bd, u_bd, ut_bd, req_bd]]

    @staticmethod
    def loss_fn(all_params, constraints):
        mu, k = all_params["static"]["problem"]["mu"], all_params["static"]["problem"]["k"]
        _, u, ut, utt = constraints[0]
        phys = mean((utt + mu * ut + k * u) ** 2)
        _, uc, utc, u, ut = constraints[1]
        if len(uc):
            return phys + 1e6 * mean((u - uc) ** 2) + 1e2 * mean((ut - utc) ** 2)
        return phys

    @staticmethod
    def residuals(all_params):
        mu, k = all_params["static"]["problem"]["mu"], all_params["static"]["problem"]["k"]
        return [Residual(0, 1.0, [(1.0, J(2), None), (mu, J(1), None), (k, J(0), None)]),
                Residual(1, 1e6, [(1.0, J(0), None), (-1.0, V(0), None)]),
                Residual(1, 1e2, [(1.0, J(1), None), (-1.0, V(1), None)])]

    @staticmethod
    def exact_solution(all_params, x_batch, batch_shape=None):
        d, w0 = all_params["static"]["problem"]["d"], all_params["static"]["problem"]["w0"]
        x = x_batch.detach().cpu().double().numpy()
        w = np.sqrt(w0 ** 2 - d ** 2)
        phi = np.arctan(-d / w)
        A = 1 / (2 * np.cos(phi))
        u = np.exp(-d * x) * 2 * A * np.cos(phi + w * x)
        return torch.tensor(u, dtype=torch.float32)


class HarmonicOscillator1DHardBC(HarmonicOscillator1D):
    """Hard-constraint ansatz u = 1 + tanh(t/sd)^2 NN (problems.py:149-206)."""

    @staticmethod
    def init_params(d=2, w0=20, sd=0.1):
        mu, k = 2 * d, w0 ** 2
        return {"dims": (1, 1), "d": d, "w0": w0, "mu": mu, "k": k, "sd": sd}, {}

    @staticmethod
    def sample_constraints(all_params, domain, key, sampler, batch_shapes):
        x_phys = domain.sample_interior(all_params, key, sampler, batch_shapes[0])
        return [[x_phys, ((0, ()), (0, (0,)), (0, (0, 0)))]]

    @staticmethod
    def constraining_fn(all_params, x_batch, u):
        sd = all_params["static"]["problem"]["sd"]
        return 1 + torch.tanh(x_batch[:, 0:1] / sd) ** 2 * u

    @staticmethod
    def loss_fn(all_params, constraints):
        mu, k = all_params["static"]["problem"]["mu"], all_params["static"]["problem"]["k"]
        _, u, ut, utt = constraints[0]
        return mean((utt + mu * ut + k * u) ** 2)

    @staticmethod
    def residuals(all_params):
        mu, k = all_params["static"]["problem"]["mu"], all_params["static"]["problem"]["k"]
        return [Residual(0, 1.0, [(1.0, J(2), None), (mu, J(1), None), (k, J(0), None)])]


class HarmonicOscillator1DInverse(HarmonicOscillator1D):
    """Learnable damping mu from 13 observations (problems.py:209-271)."""

    @staticmethod
    def init_params(d=2, w0=20):
        mu, k = 2 * d, w0 ** 2
        return ({"dims": (1, 1), "d": d, "w0": w0, "mu_true": mu, "k": k},
                {"mu": torch.tensor(0.0)})

    @staticmethod
    def sample_constraints(all_params, domain, key, sampler, batch_shapes):
        x_phys = domain.sample_interior(all_params, key, sampler, batch_shapes[0])
        req_phys = ((0, ()), (0, (0,)), (0, (0, 0)))
        x_data = torch.tensor(np.linspace(0, 1, 13).reshape((13, 1)), dtype=torch.float32)
        u_data = HarmonicOscillator1DInverse.exact_solution(all_params, x_data)
        return [[x_phys, req_phys], [x_data, u_data, ((0, ()),)]]

    @staticmethod
    def loss_fn(all_params, constraints):
        mu, k = all_params["trainable"]["problem"]["mu"], all_params["static"]["problem"]["k"]
        _, u, ut, utt = constraints[0]
        phys = mean((utt + mu * ut + k * u) ** 2)
        _, uc, u = constraints[1]
        return phys + 1e6 * mean((u - uc) ** 2)

    @staticmethod
    def residuals(all_params):
        k = all_params["static"]["problem"]["k"]
        return [Residual(0, 1.0, [(1.0, J(2), None), (lambda p: p["trainable"]["problem"]["mu"], J(1), None),
                                  (k, J(0), None)]),
                Residual(1, 1e6, [(1.0, J(0), None), (-1.0, V(0), None)])]


class BurgersEquation2D(Problem):
    """u_t + u u_x = nu u_xx on [-1,1]x[0,1], u(x,0) = -sin(pi x), zero BCs
    (problems.py:276-341)."""

    @staticmethod
    def init_params(nu=0.01 / math.pi, sd=0.1):
        return {"dims": (1, 2), "nu": nu, "sd": sd}, {}

    @staticmethod
    def sample_constraints(all_params, domain, key, sampler, batch_shapes):
        x_phys = domain.sample_interior(all_params, key, sampler, batch_shapes[0])
        return [[x_phys, ((0, ()), (0, (0,)), (0, (1,)), (0, (0, 0)))]]

    @staticmethod
    def constraining_fn(all_params, x_batch, u):
        sd = all_params["static"]["problem"]["sd"]
        x, t = x_batch[:, 0:1], x_batch[:, 1:2]
        return (torch.tanh((x + 1) / sd) * torch.tanh((1 - x) / sd) * torch.tanh((t - 0) / sd) * u
                - torch.sin(math.pi * x))

    @staticmethod
    def loss_fn(all_params, constraints):
        nu = all_params["static"]["problem"]["nu"]
        _, u, ux, ut, uxx = constraints[0]
        return mean((ut + u * ux - nu * uxx) ** 2)

    @staticmethod
    def residuals(all_params):
        nu = all_params["static"]["problem"]["nu"]
        return [Residual(0, 1.0, [(1.0, J(2), None), (1.0, J(0), J(1)), (-nu, J(3), None)])]

    @staticmethod
    def exact_solution(all_params, x_batch, batch_shape):
        nu = all_params["static"]["problem"]["nu"]
        x = x_batch.detach().cpu().double().numpy()
        vx = np.linspace(x[:, 0].min(), x[:, 0].max(), batch_shape[0])
        vt = np.linspace(x[:, 1].min(), x[:, 1].max(), batch_shape[1])
        logger.info("Running burgers_viscous_time_exact1..")
        vu = burgers_viscous_time_exact1(nu, len(vx), vx, len(vt), vt)
        return torch.tensor(vu.reshape((-1, 1)), dtype=torch.float32)


def _gaussian_sum(p, xy):
    """sum_k amp_k exp(-0.5 |xy - c_k|^2 / s_k^2) for source rows p (k, 4)."""
    p = p[:, None, :]
    return (p[:, :, 3:4] * torch.exp(-0.5 * ((xy[None] - p[:, :, 0:2]) ** 2).sum(2, keepdim=True)
                                     / (p[:, :, 2:3] ** 2))).sum(0)


class WaveEquationConstantVelocity3D(Problem):
    """(2+1)D wave equation with constant velocity and Gaussian initial
    sources, hard-constraint ansatz (problems.py:346-481)."""

    @staticmethod
    def init_params(c0=1, source=np.array([[0.0, 0.0, 0.2, 1.0]])):
        return ({"dims": (1, 3), "c0": c0, "c_fn": WaveEquationConstantVelocity3D.c_fn,
                 "source": torch.tensor(np.asarray(source), dtype=torch.float32)}, {})

    @staticmethod
    def sample_constraints(all_params, domain, key, sampler, batch_shapes):
        x_phys = domain.sample_interior(all_params, key, sampler, batch_shapes[0])
        return [[x_phys, ((0, (0, 0)), (0, (1, 1)), (0, (2, 2)))]]

    @staticmethod
    def constraining_fn(all_params, x_batch, u):
        params = all_params["static"]["problem"]
        c0, source = params["c0"], _t(params["source"], x_batch)
        xy, t = x_batch[:, 0:2], x_batch[:, 2:3]
        f = _gaussian_sum(source, xy)
        t1 = source[:, 2].min() / c0
        f = torch.exp(-0.5 * (1.5 * t / t1) ** 2) * f
        return f + torch.tanh(2.5 * t / t1) ** 2 * u

    @staticmethod
    def loss_fn(all_params, constraints):
        c_fn = all_params["static"]["problem"]["c_fn"]
        x_batch, uxx, uyy, utt = constraints[0]
        return mean(((uxx + uyy) - (1 / c_fn(all_params, x_batch) ** 2) * utt) ** 2)

    @staticmethod
    def residuals(all_params):
        def inv_c2(p, x):
            return -1 / p["static"]["problem"]["c_fn"](p, x) ** 2
        return [Residual(0, 1.0, [(1.0, J(0), None), (1.0, J(1), None), (1.0, J(2), F(inv_c2))])]

    @staticmethod
    def c_fn(all_params, x_batch):
        return torch.full((1, 1), float(all_params["static"]["problem"]["c0"]), dtype=x_batch.dtype,
                          device=x_batch.device)

    @staticmethod
    def exact_solution(all_params, x_batch, batch_shape):
        """Finite-difference reference from the CPML solver on a refined grid
        (problems.py:407-474)."""
        params = all_params["static"]["problem"]
        c0 = params["c0"]
        source = np.asarray(torch.as_tensor(params["source"]).cpu(), dtype=np.float64)
        xb = x_batch.detach().cpu().double().numpy()
        (xmin, ymin, tmin), (xmax, ymax, tmax) = xb.min(0), xb.max(0)
        deltax, deltay, deltat = [(b - a) / (n - 1) for a, b, n in zip((xmin, ymin, tmin), (xmax, ymax, tmax),
                                                                      batch_shape)]
        f0 = c0 / source[:, 2].min()
        DELTAX = DELTAY = 1 / (f0 * 10)
        DELTAT = DELTAX / (4 * np.sqrt(2) * c0)
        dx, dy, dt = int(np.ceil(deltax / DELTAX)), int(np.ceil(deltay / DELTAY)), int(np.ceil(deltat / DELTAT))
        DELTAX, DELTAY, DELTAT = deltax / dx, deltay / dy, deltat / dt
        NX, NY, NSTEPS = batch_shape[0] * dx - (dx - 1), batch_shape[1] * dy - (dy - 1), batch_shape[2] * dt - (dt - 1)
        xx, yy = np.meshgrid(np.linspace(xmin, xmax, NX), np.linspace(ymin, ymax, NY), indexing="ij")
        xy = np.stack([xx.ravel(), yy.ravel()], axis=1)
        p = source[:, None, :]
        f = (p[:, :, 3:4] * np.exp(-0.5 * ((xy[None] - p[:, :, 0:2]) ** 2).sum(2, keepdims=True)
                                   / (p[:, :, 2:3] ** 2))).sum(0)
        p0 = f.reshape((NX, NY))
        c = np.asarray(params["c_fn"](all_params, torch.tensor(xy)).cpu(), dtype=np.float64)
        c = c.reshape((NX, NY)) if c.shape[0] > 1 else c * np.ones_like(xx)
        npml = 10
        p0 = np.pad(p0, [(npml, npml), (npml, npml)], mode="edge")
        c = np.pad(c, [(npml, npml), (npml, npml)], mode="edge")
        logger.info(f"Running seismicCPML2D {(NX, NY, NSTEPS)}..")
        wavefields, _ = seismicCPML2D(NX + 2 * npml, NY + 2 * npml, NSTEPS, DELTAX, DELTAY, DELTAT, npml, c,
                                      np.ones((NX + 2 * npml, NY + 2 * npml)), (p0.copy(), p0.copy()), f0,
                                      np.float64, output_wavefields=True, gather_is=None)
        wavefields = wavefields[:, npml:-npml, npml:-npml][::dt, ::dx, ::dy]
        wavefields = np.moveaxis(wavefields, 0, -1)
        assert wavefields.shape == tuple(batch_shape)
        return torch.tensor(wavefields.reshape((-1, 1)), dtype=torch.float32)


class PoissonMultiscale2D(Problem):
    """u_xx + u_yy = sum_{k in ks} sin(2^k pi x) sin(2^k pi y) on [0,1]^2 with
    zero Dirichlet BCs through the hard ansatz
    u = tanh(x/sd) tanh((1-x)/sd) tanh(y/sd) tanh((1-y)/sd) NN.
    No reference class: written through the Problem API (problems.py:19-62).
    Exact solution u = -sum_k sin(2^k pi x) sin(2^k pi y) / (2 (2^k pi)^2)."""

    @staticmethod
    def init_params(ks=(1, 2, 3, 4), sd=0.1):
        return {"dims": (1, 2), "ks": tuple(ks), "sd": sd}, {}

    @staticmethod
    def sample_constraints(all_params, domain, key, sampler, batch_shapes):
        x_phys = domain.sample_interior(all_params, key, sampler, batch_shapes[0])
        return [[x_phys, ((0, (0, 0)), (0, (1, 1)))]]

    @staticmethod
    def constraining_fn(all_params, x_batch, u):
        sd = all_params["static"]["problem"]["sd"]
        x, y = x_batch[:, 0:1], x_batch[:, 1:2]
        return torch.tanh(x / sd) * torch.tanh((1 - x) / sd) * torch.tanh(y / sd) * torch.tanh((1 - y) / sd) * u

    @staticmethod
    def source(all_params, x_batch):
        ks = all_params["static"]["problem"]["ks"]
        x, y = x_batch[:, 0:1], x_batch[:, 1:2]
        return sum(torch.sin(2 ** k * math.pi * x) * torch.sin(2 ** k * math.pi * y) for k in ks)

    @staticmethod
    def loss_fn(all_params, constraints):
        x, uxx, uyy = constraints[0]
        return mean((uxx + uyy - PoissonMultiscale2D.source(all_params, x)) ** 2)

    @staticmethod
    def residuals(all_params):
        return [Residual(0, 1.0, [(1.0, J(0), None), (1.0, J(1), None), (-1.0, F(PoissonMultiscale2D.source), None)])]

    @staticmethod
    def exact_solution(all_params, x_batch, batch_shape=None):
        ks = all_params["static"]["problem"]["ks"]
        x = x_batch.detach().cpu().double().numpy()
        u = np.zeros((x.shape[0], 1))
        for k in ks:
            a = 2 ** k * np.pi
            u[:, 0] -= np.sin(a * x[:, 0]) * np.sin(a * x[:, 1]) / (2 * a * a)
        return torch.tensor(u, dtype=torch.float32)


def wave_sources(seed=2, sigmas=(0.3, 0.1, 0.03), amp=1.0, lo=-0.5, hi=0.5):
    """Sources (k, 4) rows [cx, cy, sigma, amp] with centres drawn from
    U(lo, hi)^2 by the seeded key."""
    c = rng.uniform(rng.PRNGKey(seed), (len(sigmas), 2), lo, hi)
    return np.concatenate([c, np.asarray(sigmas)[:, None], amp * np.ones((len(sigmas), 1))], axis=1)


class BatchedHarmonicOscillatorInverse(Problem):
    """n_inst independent inverse damped oscillators u'' + mu_i u' + k_i u = 0,
    u(0) = 1, u'(0) = 0, each with its own learnable mu_i (init 0) fitted to
    n_obs noisy observations of its analytic solution, as in
    HarmonicOscillator1DInverse (problems.py:209-271) with a per-instance
    coefficient.  Instance i lives on x in [spacing*i, spacing*i + 1]
    (InstancesDomain1D); the loss is the sum of the instances' losses, so each
    instance's parameters receive exactly their own gradient."""

    @staticmethod
    def init_params(w0s, mus_true, n_obs=64, noise=0.01, seed=3, spacing=2.0, data_weight=1e6):
        w0s = np.asarray(w0s, dtype=np.float64)
        mus_true = np.asarray(mus_true, dtype=np.float64)
        n = len(w0s)
        static = {"dims": (1, 1), "n_inst": n, "spacing": float(spacing), "w0": w0s, "mu_true": mus_true,
                  "d": mus_true / 2, "k": torch.tensor(w0s ** 2, dtype=torch.float32), "n_obs": int(n_obs),
                  "noise": float(noise), "seed": int(seed), "data_weight": float(data_weight)}
        return static, {"mu": torch.zeros(n, dtype=torch.float32)}

    @staticmethod
    def _exact_local(d, w0, t):
        w = np.sqrt(w0 ** 2 - d ** 2)
        phi = np.arctan(-d / w)
        return np.exp(-d * t) * 2 * (1 / (2 * np.cos(phi))) * np.cos(phi + w * t)

    @staticmethod
    def sample_constraints(all_params, domain, key, sampler, batch_shapes):
        p = all_params["static"]["problem"]
        n, sp = p["n_inst"], p["spacing"]
        x_phys = domain.sample_interior(all_params, key, sampler, batch_shapes[0])
        inst_phys = torch.repeat_interleave(torch.arange(n), x_phys.shape[0] // n)
        t_obs = np.linspace(0.0, 1.0, p["n_obs"])
        x_obs = (sp * np.arange(n)[:, None] + t_obs[None]).reshape(-1, 1)
        u = BatchedHarmonicOscillatorInverse._exact_local(p["d"][:, None], p["w0"][:, None], t_obs[None])
        u = u + p["noise"] * rng.normal(rng.PRNGKey(p["seed"]), u.shape)
        inst_obs = torch.repeat_interleave(torch.arange(n), p["n_obs"])
        return [[x_phys, inst_phys, ((0, ()), (0, (0,)), (0, (0, 0)))],
                [torch.tensor(x_obs, dtype=torch.float32), torch.tensor(u.reshape(-1, 1), dtype=torch.float32),
                 inst_obs, ((0, ()),)]]

    @staticmethod
    def loss_fn(all_params, constraints):
        p = all_params["static"]["problem"]
        mu = all_params["trainable"]["problem"]["mu"]
        n = p["n_inst"]
        _, inst, u, ut, utt = constraints[0]
        k = p["k"].to(u.device)
        r = utt + mu[inst][:, None] * ut + k[inst][:, None] * u
        _, uc, inst_o, u_o = constraints[1]
        # sum over instances of per-instance means (equal point counts per instance)
        return n * mean(r ** 2) + p["data_weight"] * n * mean((u_o - uc) ** 2)

    @staticmethod
    def residuals(all_params):
        p = all_params["static"]["problem"]
        n = p["n_inst"]

        def k_inst(ap, x, inst):
            return torch.as_tensor(ap["static"]["problem"]["k"]).to(x.device, x.dtype)[inst.long()][:, None]
        return [Residual(0, float(n), [(1.0, J(2), None),
                                       (1.0, P(lambda ap: ap["trainable"]["problem"]["mu"], 0), J(1)),
                                       (1.0, F(k_inst), J(0))]),
                Residual(1, p["data_weight"] * n, [(1.0, J(0), None), (-1.0, V(0), None)])]

    @staticmethod
    def exact_solution(all_params, x_batch, batch_shape=None):
        p = all_params["static"]["problem"]
        x = x_batch.detach().cpu().double().numpy()[:, 0]
        inst = np.floor(x / p["spacing"] + 1e-9).astype(int)
        t = x - p["spacing"] * inst
        u = BatchedHarmonicOscillatorInverse._exact_local(p["d"][inst], p["w0"][inst], t)
        return torch.tensor(u[:, None], dtype=torch.float32)
