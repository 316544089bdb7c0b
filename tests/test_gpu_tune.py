"""GPU: the autotuner (libxstream_amd/tune_multiply) end to end on one small shape.

It writes a CSV row in the MI355X format (with the SPLIT/BSS/SPLITS columns), and that row loads
through libsmm_acc_params_load and gives the oracle's C through libsmm_acc_process (the parity bar).
"""
import os
import subprocess

import numpy as np
import pytest

import libxstream_amd as lx
from libxstream_amd.workload import make_stack_workload
from tests.helpers import oracle_error, run_stack

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_tune_writes_loadable_row(tmp_path):
    csv = tmp_path / "tuned.csv"
    exe = os.path.join(ROOT, "libxstream_amd", "tune_multiply")
    run = subprocess.run([exe, "-t", "f64", "-s", "3000", "-r", "3", "-o", str(csv), "13", "13", "13"],
                         capture_output=True, text=True, timeout=240)
    assert run.returncode == 0, run.stdout + run.stderr
    rows = csv.read_text().splitlines()
    assert rows[0] == "DEVICE;TYPEID;M;N;K;S;GFLOPS;VARIANT;TM;TN;NT;BS;KT;FLUSH;SPLIT;BSS;SPLITS"
    fields = rows[1].split(";")
    assert len(fields) == 17 and fields[1:6] == ["3", "13", "13", "13", "3000"] and float(fields[6]) > 0
    lx.init()
    s = lx.Stream("tune")
    try:
        # a registered row without GFLOPS loses to any measured duplicate (smm_acc.c:117-119)
        lx.config_set(lx.REAL_8, 13, 13, 13, lx.SmmConfig(variant=1, tm=1, tn=1, nt=64, bs=16, kt=16, flush=2))
        assert lx.lib().libsmm_acc_params_load(str(csv).encode()) == 1
        cfg, src = lx.config_get(lx.REAL_8, 13, 13, 13)
        assert src == 1 and cfg.s == 3000 and cfg.bs == int(fields[11]) and cfg.split == int(fields[14])
        for size in (3000, 500):  # at S (the short-stack pair, if any) and below it (bs scaled down)
            w = make_stack_workload(13, 13, 13, size, dtype=np.float64, seed=size, c_init="uniform")
            assert oracle_error(w, run_stack(s, w))[0] <= 1e-13
    finally:  # back to the shipped row
        lx.config_set(lx.REAL_8, 13, 13, 13, lx.SmmConfig(variant=1, tm=1, tn=1, nt=64, bs=16, kt=16, flush=2))
        lx.lib().libsmm_acc_params_load(os.path.join(ROOT, "libxstream_amd", "params",
                                                     "tune_multiply_MI355X.csv").encode())
        s.close()
        lx.finalize()
