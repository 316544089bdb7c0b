"""Host-side checks of the flush contract (no GPU needed).

The stack kernels write a C-run either with atomic adds, which is correct for any stack order
(samples/smm/kernels/multiply.cl:657,687), or with a plain read-modify-write that is only correct when
each C block is one contiguous run of the stack. Configurations that libsmm_acc_process reaches by itself
(the shipped CSV, the learned predictor, the built-in default) must therefore not select the
read-modify-write flush unless LIBXSTREAM_SMM_GROUPED declares the stacks grouped.
"""
import csv
import os

import numpy as np
import pytest

import libxstream_amd as lx
import libxstream_amd.learned as learned

PARAMS = os.path.join(os.path.dirname(lx.LIB_PATH), "params")


@pytest.fixture(scope="module")
def lib():
    lx.init()
    yield lx
    lx.finalize()


def test_shipped_csv_has_no_rmw_rows():
    with open(os.path.join(PARAMS, "tune_multiply_MI355X.csv")) as f:
        rows = list(csv.DictReader(f, delimiter=";"))
    assert rows and all(int(r["FLUSH"]) != 1 for r in rows)


def test_predictor_never_picks_rmw(lib):
    if os.environ.get("LIBXSTREAM_SMM_GROUPED", "0") not in ("", "0"):
        pytest.skip("stacks declared grouped")
    model = lx.Mlp(path=os.path.join(PARAMS, "mlp_MI355X.txt"))
    try:
        rng = np.random.default_rng(7)
        for _ in range(24):
            tid = (lx.REAL_8, lx.REAL_4)[int(rng.integers(0, 2))]
            m, n, k = (int(v) for v in rng.integers(4, 81, 3))
            cfg, _ = model.predict_config(tid, m, n, k)
            assert cfg.flush == 2, (tid, m, n, k, cfg)
            assert all(c.flush == 2 for c, _ in model.predict_topk(tid, m, n, k, 8))
    finally:
        model.close()


def test_resolved_configs_never_rmw(lib):
    """config_get over tuned, learned and default shapes"""
    if os.environ.get("LIBXSTREAM_SMM_GROUPED", "0") not in ("", "0"):
        pytest.skip("stacks declared grouped")
    for tid in (lx.REAL_8, lx.REAL_4):
        for shape in ((23, 23, 23), (45, 45, 13), (5, 5, 5), (30, 60, 67), (79, 24, 36), (80, 80, 80), (4, 4, 4)):
            cfg, src = lx.config_get(tid, *shape)
            # flush 1 in a resolved configuration would launch the read-modify-write flush
            assert cfg.flush != 1, (tid, shape, src, cfg)


def test_heldout_scoring_uses_runnable_rows():
    rows = [{"variant": 1, "tm": 2, "tn": 2, "nt": 64, "kt": 24, "bs": 16, "flush": 1, "gflops": 10.0},
            {"variant": 1, "tm": 2, "tn": 2, "nt": 64, "kt": 24, "bs": 16, "flush": 2, "gflops": 9.0}]
    if learned.grouped():
        pytest.skip("stacks declared grouped")
    assert learned.best_row(learned.runnable(rows))["gflops"] == 9.0
