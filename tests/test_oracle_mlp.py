"""Pins of the oracle's predictor-network restatement (oracle/mlp_oracle.c), CPU only.

The analytic gradient is fixed by central finite differences of the loss on a network with the
predictor's head layout (softmax heads with soft targets, one regression head), by the closed form
of a one-layer linear regression, and planted mistakes (no ReLU mask in the backward pass, no bias
gradient) must fail the finite-difference pin.
"""
import numpy as np
import pytest

import oracle

# head layout of the predictor: variant 2, tm 5, tn 5, nt 6, kt 5, bs 4, flush 2, log2 GFLOP/s 1
SEG = np.array([0, 2, 7, 12, 18, 23, 27, 29, 30], dtype=np.int32)


def _net(rng, width):
    params = []
    for nin, nout in zip(width[:-1], width[1:]):
        params.append(rng.uniform(-1, 1, nout * nin) * np.sqrt(3.0 / nin))
        params.append(rng.uniform(-0.1, 0.1, nout))
    return np.concatenate(params)


def _targets(rng, batch):
    t = np.zeros((batch, SEG[-1]))
    for s in range(len(SEG) - 1):
        c0, c1 = SEG[s], SEG[s + 1]
        if c1 - c0 == 1:
            t[:, c0] = rng.uniform(5, 15, batch)
        else:
            raw = rng.uniform(0, 1, (batch, c1 - c0)) ** 4
            t[:, c0:c1] = raw / raw.sum(axis=1, keepdims=True)
    t[0, SEG[2]:SEG[3]] = 0.0  # an unlabelled head contributes nothing
    return t


def _fd(width, params, x, t, hw, h=1e-5):
    g = np.empty_like(params)
    for i in range(params.size):
        p = params.copy()
        p[i] += h
        lp, _ = oracle.mlp_loss_grad(width, p, x, t, SEG, hw)
        p[i] -= 2 * h
        lm, _ = oracle.mlp_loss_grad(width, p, x, t, SEG, hw)
        g[i] = (lp - lm) / (2 * h)
    return g


def _case(seed=1):
    rng = np.random.default_rng(seed)
    width = [8, 12, 10, int(SEG[-1])]
    params = _net(rng, width)
    x = rng.uniform(-1.5, 1.5, (6, 8))
    t = _targets(rng, 6)
    hw = rng.uniform(0.5, 2.0, len(SEG) - 1)
    return width, params, x, t, hw


def test_gradient_matches_finite_differences():
    width, params, x, t, hw = _case()
    loss, g = oracle.mlp_loss_grad(width, params, x, t, SEG, hw)
    fd = _fd(width, params, x, t, hw)
    assert loss > 0
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-7)  # loss ~ 1e2: cancellation ~ 1e2 * 2^-53 / h


def test_linear_regression_closed_form():
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (5, 3))
    w, b = rng.uniform(-1, 1, 3), 0.3
    t = rng.uniform(-1, 1, (5, 1))
    seg = np.array([0, 1], dtype=np.int32)
    hw = np.array([1.7])
    loss, g = oracle.mlp_loss_grad([3, 1], np.concatenate([w, [b]]), x, t, seg, hw)
    r = x @ w + b - t[:, 0]
    assert np.isclose(loss, 1.7 * np.mean(r ** 2), rtol=1e-14)
    assert np.allclose(g[:3], 2 * 1.7 * (r[:, None] * x).mean(axis=0), rtol=1e-13)
    assert np.isclose(g[3], 2 * 1.7 * r.mean(), rtol=1e-13)


@pytest.mark.parametrize("plant", [oracle.MLP_PLANT_NORELU, oracle.MLP_PLANT_NOBIAS])
def test_planted_mistakes_are_caught(plant):
    width, params, x, t, hw = _case(3)
    _, g = oracle.mlp_loss_grad(width, params, x, t, SEG, hw, plant=plant)
    assert not np.allclose(g, _fd(width, params, x, t, hw), rtol=1e-5, atol=1e-7)
