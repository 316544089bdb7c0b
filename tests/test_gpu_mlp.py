"""GPU checks of the learned predictor (libsmm_acc_mlp_*, HIP kernels in mlp.hip).

- Gradients of the HIP backward pass match the double-precision oracle (oracle/mlp_oracle.c, pinned
  by finite differences), for the predictor's own layout (41 inputs, 3 x 64 ReLU, 2 outputs).
- Full-batch Adam on the GPU reduces the loss of a learnable synthetic task.
- Predicted configurations are valid for their shapes; an installed model serves shapes without a
  tuned configuration, and a threshold it cannot reach leaves the default in place.
- The shipped model (params/mlp_MI355X.txt), trained on the shipped sweeps' 400 training triplets,
  is judged on the 100 held-out triplets by the exhaustive autotuner's measurements: it must reach
  90% of the best on average and beat both the default heuristic and the reference's
  nearest-tuned-triplet method; a few of its choices are also timed live against the best.
"""
import os

import numpy as np
import pytest

import libxstream_amd as lx
import libxstream_amd.learned as learned
import oracle
from libxstream_amd.workload import make_stack_workload

pytestmark = pytest.mark.gpu

PARAMS = os.path.join(os.path.dirname(lx.LIB_PATH), "params")
MODEL = os.path.join(PARAMS, "mlp_MI355X.txt")
SWEEPS = [os.path.join(PARAMS, "sweep_f64_MI355X.jsonl.gz"), os.path.join(PARAMS, "sweep_f32_MI355X.jsonl.gz")]


@pytest.fixture(scope="module")
def dev():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    yield
    lx.finalize()


def _inputs(rng, n):
    """n rows of real (shape, configuration) inputs"""
    x = []
    while len(x) < n:
        tid = [lx.REAL_8, lx.REAL_4][int(rng.integers(0, 2))]
        m, k, nn = (int(v) for v in rng.integers(4, 81, 3))
        cfg = lx.default_config(tid, m, nn, k)
        cfg.bs = int(rng.choice([16, 32, 64, 128]))
        cfg.flush = int(rng.integers(1, 3))
        x.append(lx.mlp_features(tid, m, nn, k, cfg))
    return np.array(x, dtype=np.float32)


def test_gradient_matches_oracle(dev):
    rng = np.random.default_rng(1)
    model = lx.Mlp(1, 64, 3, seed=7)
    try:
        info = model.info()
        assert info["width"] == [41, 64, 64, 64, 2]
        x = _inputs(rng, 96)
        t = np.stack([rng.uniform(0.3, 1.0, 96), rng.uniform(8, 15, 96)], axis=1).astype(np.float32)
        w = np.array(learned.OUT_WEIGHTS, dtype=np.float32)
        loss, g = model.gradient(x, t, w)
        ref_loss, ref_g = oracle.mlp_loss_grad(info["width"], model.params().astype(np.float64), x, t,
                                               np.array([0, 1, 2], dtype=np.int32), w)
    finally:
        model.close()
    assert np.isclose(loss, ref_loss, rtol=1e-5)
    assert np.abs(g - ref_g).max() <= 1e-4 * np.abs(ref_g).max()  # FP32 against FP64
    assert np.corrcoef(g, ref_g)[0, 1] > 0.99999


def test_host_inference_matches_device(dev):
    """Prediction runs the ensemble on the host (no device work on the launch path); it performs the same
    operations in the same order as the device forward pass, so the outputs agree to the last bit or
    within one rounding of FP32 where the compiler chose a different but equivalent instruction."""
    rng = np.random.default_rng(9)
    x = _inputs(rng, 300)
    model = lx.Mlp(3, 64, 3, seed=21)
    try:
        dev_out = model.forward(x)
        host_out = model.forward_host(x)
    finally:
        model.close()
    assert np.allclose(host_out, dev_out, rtol=2e-6, atol=1e-6)
    assert np.mean(host_out == dev_out) > 0.95


def test_training_reduces_loss(dev):
    rng = np.random.default_rng(2)
    x = _inputs(rng, 512)
    t = np.stack([0.5 + 0.3 * np.tanh(x[:, 0] - x[:, 2] + x[:, 9]), 10 + 2 * x[:, 7]], axis=1).astype(np.float32)
    model = lx.Mlp(2, 64, 3, seed=3)
    try:
        loss0, _ = model.gradient(x, t)
        loss1 = model.train(x, t, epochs=300, lr=3e-3)
        loss2 = model.train(x, t, epochs=300, lr=1e-3)
    finally:
        model.close()
    assert np.isfinite(loss2) and loss1 < loss0 / 10 and loss2 <= loss1 * 1.05


@pytest.mark.parametrize("tid", [lx.REAL_8, lx.REAL_4])
def test_predicted_configs_valid_and_installed(dev, tid, tmp_path):
    model = lx.Mlp(2, 64, 3, seed=11)
    try:
        for (m, n, k) in ((4, 4, 4), (23, 23, 23), (7, 60, 33), (80, 80, 80), (45, 45, 13)):
            cfg, conf = model.predict_config(tid, m, n, k)
            assert lx.config_valid(tid, m, n, k, cfg) and 0 <= conf <= 1
        path = str(tmp_path / "mlp.txt")
        model.save(path)
        again = lx.Mlp(path=path)
        try:
            assert all(np.array_equal(again.params(i), model.params(i)) for i in range(2))
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


@pytest.mark.parametrize("tid", [lx.REAL_8, lx.REAL_4])
def test_predict_topk(dev, tid):
    """the ranking behind learned-guided tuning: best first, distinct, valid, led by predict_config's choice"""
    model = lx.Mlp(2, 64, 3, seed=5)
    try:
        for (m, n, k) in ((4, 4, 4), (23, 23, 23), (7, 60, 33), (45, 45, 13)):
            ranked = model.predict_topk(tid, m, n, k, 8)
            assert 1 <= len(ranked) <= 8
            scores = [sc for _, sc in ranked]
            assert scores == sorted(scores, reverse=True)
            keys = {learned.config_key(c) for c, _ in ranked}
            assert len(keys) == len(ranked)
            assert all(lx.config_valid(tid, m, n, k, c) for c, _ in ranked)
            top, conf = model.predict_config(tid, m, n, k)
            assert learned.config_key(top) == learned.config_key(ranked[0][0])
            assert conf == pytest.approx(min(max(scores[0], 0.0), 1.0))
            assert len(model.predict_topk(tid, m, n, k, 10000)) >= len(ranked)
    finally:
        model.close()


@pytest.fixture(scope="module")
def shipped(dev):
    if not os.path.exists(MODEL):
        pytest.skip("no shipped model")
    sweep = learned.load_sweep(SWEEPS)
    train_t, held_t = learned.split(learned.triplets())
    tids = (lx.REAL_8, lx.REAL_4)
    train_keys = [(tid,) + t for tid in tids for t in train_t if (tid,) + t in sweep]
    held_keys = [(tid,) + t for tid in tids for t in held_t if (tid,) + t in sweep]
    model = lx.Mlp(path=MODEL)
    yield model, sweep, train_keys, held_keys
    model.close()


def test_shipped_model_on_heldout(shipped):
    model, sweep, train_keys, held_keys = shipped
    assert len(held_keys) == 200
    results = learned.evaluate(model, sweep, held_keys, train_keys)
    rep = learned.report(results, (lx.REAL_8, lx.REAL_4))
    print("held-out: learned %s default %s nearest %s" % (rep["learned"], rep["default"], rep["nearest"]))
    assert all(r["learned"] is not None for r in results)  # every choice was a measured configuration
    assert rep["learned"]["mean"] >= 0.9
    # per triplet (a lookup in the shipped sweeps, so deterministic): at least 80% of the held-out choices
    # within 10% of the sweep's best measured configuration, 95% for the learned-guided top 4
    assert rep["learned"]["within_10pct"] >= 0.8 and rep["learned_top4"]["within_10pct"] >= 0.95
    assert rep["learned"]["mean"] > rep["default"]["mean"] and rep["learned"]["mean"] > rep["nearest"]["mean"]
    assert rep["type3"]["learned"]["mean"] >= 0.9
    # learned-guided tuning: measuring more of the network's ranking never loses
    for f in ("learned_top2", "learned_top4", "learned_top8"):
        assert all(r[f] is not None for r in results)
    assert rep["learned"]["mean"] <= rep["learned_top2"]["mean"] <= rep["learned_top4"]["mean"] <= rep["learned_top8"]["mean"]


def test_shipped_model_live(shipped):
    """Held-out choices timed now (learned.evaluate_live): the network's choice, its top 4 and the sweep's
    3 fastest configurations timed together on a fresh 30000-entry stack, interleaved rounds, medians. Per
    triplet: at least 80% of the network's choices within 10% of the fastest of the sweep's 3, and every
    learned-guided top-4 result within 30%."""
    model, sweep, _, held_keys = shipped
    keys = held_keys[:100:5] + held_keys[100::5]  # 20 FP64 and 20 FP32 held-out triplets
    stream = lx.Stream("live")
    try:
        results = learned.evaluate_live(model, sweep, keys, stream=stream)
    finally:
        stream.close()
    rep = {f: learned.summary_live(results, f) for f in ("learned", "learned_top4")}
    print("live held-out:", rep)
    assert rep["learned"]["within_10pct"] >= 0.8
    assert rep["learned_top4"]["min"] >= 0.7 and rep["learned_top4"]["within_10pct"] >= 0.9
