"""GPU checks of the learned predictor's network (libsmm_acc_mlp_*, HIP kernels in mlp.hip).

Gradients of the HIP backward pass are compared with the double-precision oracle
(oracle/mlp_oracle.c, itself pinned by finite differences); full-batch Adam training on the GPU
must reduce the loss of a learnable synthetic task; predicted configurations must be valid for
their shapes, and an installed model must serve shapes that have no tuned configuration.
"""
import numpy as np
import pytest

import libxstream_amd as lx
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    yield
    lx.finalize()


def _soft_targets(rng, batch, seg):
    t = np.zeros((batch, seg[-1]), dtype=np.float32)
    for s in range(len(seg) - 1):
        c0, c1 = seg[s], seg[s + 1]
        if c1 - c0 == 1:
            t[:, c0] = rng.uniform(5, 15, batch)
        else:
            raw = rng.uniform(0, 1, (batch, c1 - c0)) ** 4
            t[:, c0:c1] = raw / raw.sum(axis=1, keepdims=True)
    return t


def _width():
    seg, _ = lx.mlp_heads()
    return [8, 64, 64, 64, int(seg[-1])], seg


def test_gradient_matches_oracle(dev):
    width, seg = _width()
    rng = np.random.default_rng(1)
    model = lx.Mlp(64, 3, seed=7)
    try:
        x = rng.uniform(-1.5, 1.5, (96, 8)).astype(np.float32)
        t = _soft_targets(rng, 96, seg)
        hw = rng.uniform(0.5, 2, len(seg) - 1).astype(np.float32)
        loss, g = model.gradient(x, t, hw)
        ref_loss, ref_g = oracle.mlp_loss_grad(width, model.params().astype(np.float64), x, t, seg, hw)
    finally:
        model.close()
    assert np.isclose(loss, ref_loss, rtol=1e-5)
    scale = np.abs(ref_g).max()
    assert np.abs(g - ref_g).max() <= 1e-4 * scale  # FP32 against FP64, K <= 96 sums
    assert np.corrcoef(g, ref_g)[0, 1] > 0.99999


def test_training_reduces_loss(dev):
    _, seg = _width()
    rng = np.random.default_rng(2)
    n = 256
    x = rng.uniform(-1.5, 1.5, (n, 8)).astype(np.float32)
    t = np.zeros((n, seg[-1]), dtype=np.float32)
    # a learnable rule: each head's class is a function of the sign pattern of two features
    for h in range(len(seg) - 2):
        c0, c1 = seg[h], seg[h + 1]
        cls = ((x[:, h % 8] > 0).astype(int) + 2 * (x[:, (h + 3) % 8] > 0)) % (c1 - c0)
        t[np.arange(n), c0 + cls] = 1.0
    t[:, seg[-2]] = 8 + 2 * x[:, 7]
    model = lx.Mlp(64, 3, seed=3)
    try:
        loss0, _ = model.gradient(x, t)
        loss1 = model.train(x, t, epochs=400, lr=3e-3)
        loss2 = model.train(x, t, epochs=400, lr=1e-3)
    finally:
        model.close()
    assert np.isfinite(loss2) and loss1 < loss0 / 3 and loss2 <= loss1 * 1.05


@pytest.mark.parametrize("tid", [lx.REAL_8, lx.REAL_4])
def test_predicted_configs_valid_and_installed(dev, tid, tmp_path):
    model = lx.Mlp(64, 3, seed=11)
    try:
        for (m, n, k) in ((4, 4, 4), (23, 23, 23), (7, 60, 33), (80, 80, 80), (45, 45, 13)):
            cfg, conf = model.predict_config(tid, m, n, k)
            assert lx.config_valid(tid, m, n, k, cfg) and 0 < conf <= 1
        path = str(tmp_path / "mlp.txt")
        model.save(path)
        again = lx.Mlp(path=path)
        try:
            assert np.array_equal(again.params(), model.params())
            a, _ = again.predict_config(tid, 19, 37, 53)
            b, _ = model.predict_config(tid, 19, 37, 53)
            assert a.as_dict() == b.as_dict()
        finally:
            again.close()
        model.install(0.0)
        cfg, src = lx.config_get(tid, 61, 17, 29)  # no tuned row for this shape
        assert src == 3 and lx.config_valid(tid, 61, 17, 29, cfg)
        model.install(1.01)  # no prediction can be that confident: the default stays
        cfg, src = lx.config_get(tid, 62, 17, 29)
        assert src == 0
    finally:
        lx.learned_clear()
        model.close()
