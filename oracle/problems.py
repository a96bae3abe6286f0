"""Oracle: constraining_fn / loss_fn of the config problems (torch, any dtype).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Each factory returns (required_ujs per constraint, constraining_fn(x, u),
loss_fn(constraints_, *trainable)) restating:
  * HarmonicOscillator1D           problems.py:68-132 (two constraints)
  * HarmonicOscillator1DHardBC     problems.py:149-206
  * HarmonicOscillator1DInverse    problems.py:209-271 (trainable mu)
  * BurgersEquation2D              problems.py:276-326
  * WaveEquationConstantVelocity3D problems.py:346-405
  * Poisson2D multiscale: no reference class; written through the
    reference's Problem API (problems.py:19-62).  Ansatz
    u = tanh(x/sd) tanh((1-x)/sd) tanh(y/sd) tanh((1-y)/sd) * NN, residual
    u_xx + u_yy - sum_k sin(2^k pi x) sin(2^k pi y).

Pinned in tests/test_oracle_problems.py: every loss_fn vanishes on its
problem's exact solution (oracle/exact.py, derivatives by float64 central
differences) and rejects the same fields with one derivative scaled by
1 + 1e-3; the inverse oscillator's trainable-coefficient and 1e6 data terms
and the soft boundary terms match their closed forms; each hard-constraint
ansatz, differentiated by the nested jvp for an arbitrary network output,
holds its boundary / initial values and derivatives exactly and has the
hand-expanded Taylor coefficients at t = 0 (which fix sd and t1); the Poisson
source matches hand values.
"""

import math

import torch


def ho_hard(d=2, w0=80, sd=0.1):
    mu, k = 2 * d, w0 ** 2
    req = [((0, ()), (0, (0,)), (0, (0, 0)))]

    def constraining_fn(x, u):
        return 1 + torch.tanh(x[:, 0:1] / sd) ** 2 * u

    def loss_fn(cs):
        _, u, ut, utt = cs[0]
        return torch.mean((utt + mu * ut + k * u) ** 2)
    return req, constraining_fn, loss_fn


def ho_soft(d=2, w0=20):
    mu, k = 2 * d, w0 ** 2
    req = [((0, ()), (0, (0,)), (0, (0, 0))), ((0, ()), (0, (0,)))]

    def loss_fn(cs):
        _, u, ut, utt = cs[0]
        phys = torch.mean((utt + mu * ut + k * u) ** 2)
        _, uc, utc, u, ut = cs[1]
        return phys + 1e6 * torch.mean((u - uc) ** 2) + 1e2 * torch.mean((ut - utc) ** 2)
    return req, (lambda x, u: u), loss_fn


def ho_inverse(k):
    req = [((0, ()), (0, (0,)), (0, (0, 0))), ((0, ()),)]

    def loss_fn(cs, mu):
        _, u, ut, utt = cs[0]
        phys = torch.mean((utt + mu * ut + k * u) ** 2)
        _, uc, u = cs[1]
        return phys + 1e6 * torch.mean((u - uc) ** 2)
    return req, (lambda x, u: u), loss_fn


def burgers(nu=0.01 / math.pi, sd=0.1):
    req = [((0, ()), (0, (0,)), (0, (1,)), (0, (0, 0)))]

    def constraining_fn(x, u):
        xx, t = x[:, 0:1], x[:, 1:2]
        return (torch.tanh((xx + 1) / sd) * torch.tanh((1 - xx) / sd) * torch.tanh((t - 0) / sd) * u
                - torch.sin(math.pi * xx))

    def loss_fn(cs):
        _, u, ux, ut, uxx = cs[0]
        return torch.mean((ut + u * ux - nu * uxx) ** 2)
    return req, constraining_fn, loss_fn


def wave(source, c0=1.0):
    source = torch.as_tensor(source)
    req = [((0, (0, 0)), (0, (1, 1)), (0, (2, 2)))]

    def constraining_fn(x, u):
        s = source.to(x.dtype)
        xy, t = x[:, 0:2], x[:, 2:3]
        p = s[:, None, :]
        f = (p[:, :, 3:4] * torch.exp(-0.5 * ((xy[None] - p[:, :, 0:2]) ** 2).sum(2, keepdim=True)
                                      / (p[:, :, 2:3] ** 2))).sum(0)
        t1 = s[:, 2].min() / c0
        f = torch.exp(-0.5 * (1.5 * t / t1) ** 2) * f
        return f + torch.tanh(2.5 * t / t1) ** 2 * u

    def loss_fn(cs):
        _, uxx, uyy, utt = cs[0]
        return torch.mean(((uxx + uyy) - (1 / c0 ** 2) * utt) ** 2)
    return req, constraining_fn, loss_fn


def poisson_source(x, ks=(1, 2, 3, 4)):
    return sum(torch.sin(2 ** k * math.pi * x[:, 0:1]) * torch.sin(2 ** k * math.pi * x[:, 1:2]) for k in ks)


def poisson(sd=0.1, ks=(1, 2, 3, 4)):
    req = [((0, (0, 0)), (0, (1, 1)))]

    def constraining_fn(x, u):
        xx, yy = x[:, 0:1], x[:, 1:2]
        return (torch.tanh(xx / sd) * torch.tanh((1 - xx) / sd)
                * torch.tanh(yy / sd) * torch.tanh((1 - yy) / sd) * u)

    def loss_fn(cs):
        x, uxx, uyy = cs[0]
        return torch.mean((uxx + uyy - poisson_source(x, ks)) ** 2)
    return req, constraining_fn, loss_fn
