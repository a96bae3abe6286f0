"""Pins oracle/adam.py: Kingma & Ba Algorithm 1 closed forms (the first step
moves every coordinate by lr * sign(g) up to eps; a constant gradient moves
it by exactly lr * g / (|g| + eps) per step) and an independent scalar loop
over many steps (bias corrections at counts > 1), in float64."""

import math

import numpy as np

from oracle import adam as oadam


def test_adam_first_step_is_signed_lr():
    r = np.random.default_rng(0)
    p = r.standard_normal(50)
    g = r.standard_normal(50)
    p1, m, v, c = oadam.adam_update(p, np.zeros(50), np.zeros(50), g, 0, lr=1e-3, eps=1e-8)
    assert c == 1
    assert np.allclose(p1, p - 1e-3 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-15)


def test_adam_constant_gradient_closed_form():
    g = np.array([0.5, -2.0, 1e-3])
    p = np.zeros(3)
    m = np.zeros(3)
    v = np.zeros(3)
    c = 0
    for k in range(1, 20):
        p, m, v, c = oadam.adam_update(p, m, v, g, c, lr=0.01)
        assert np.allclose(p, -0.01 * k * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)


def test_adam_matches_scalar_loop():
    r = np.random.default_rng(1)
    gs = r.standard_normal((30, 4))
    p = r.standard_normal(4)
    m = np.zeros(4)
    v = np.zeros(4)
    c = 0
    ps = p.copy()
    for t, g in enumerate(gs, start=1):
        p, m, v, c = oadam.adam_update(p, m, v, g, c, lr=3e-3, b1=0.8, b2=0.99, eps=1e-7)
    # independent scalar loop (Kingma & Ba, Algorithm 1)
    for j in range(4):
        mj = vj = 0.0
        pj = ps[j]
        for t, g in enumerate(gs[:, j], start=1):
            mj = 0.8 * mj + 0.2 * g
            vj = 0.99 * vj + 0.01 * g * g
            pj -= 3e-3 * (mj / (1 - 0.8 ** t)) / (math.sqrt(vj / (1 - 0.99 ** t)) + 1e-7)
        assert abs(pj - p[j]) < 1e-12
