"""acc_bench (libxstream_amd/acc_bench, acc_bench_f32): the reference's DBCSR/ACC stack benchmark
(samples/smm/acc_bench.c) on this library. Its report must carry the reference's lines, and its own host
check must find the device result within the parity bar."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "libxstream_amd", "acc_bench")


def _run(binary, args, env=None):
    e = dict(os.environ)
    e.update(env or {})
    p = subprocess.run([binary] + [str(a) for a in args], capture_output=True, text=True, timeout=300, env=e)
    return p.returncode, p.stdout, p.stderr


def test_built_and_reports_missing_device():
    assert os.access(BIN, os.X_OK) and os.access(BIN + "_f32", os.X_OK)
    rc, out, err = _run(BIN, [1, 100])
    if 0 == rc and "Activated device" in out:
        pytest.skip("a device is visible")
    assert rc != 0 and "No ACC-device found" in err


@pytest.mark.gpu
@pytest.mark.parametrize("binary,bar", [("acc_bench", 1e-13), ("acc_bench_f32", 1e-5)])
@pytest.mark.parametrize("args", [(3, 3000, 23), (2, 3000, 4, 9, 22), (2, 2000, 5), (2, 1000, 64)],
                         ids=["23", "4x9x22", "5", "64"])
def test_report_and_host_check(binary, bar, args):
    rc, out, err = _run(os.path.join(ROOT, "libxstream_amd", binary), args, env={"CHECK": "1"})
    assert rc == 0, out + err
    lines = out.splitlines()
    m = args[2]
    n = args[3] if len(args) > 3 else m
    k = args[4] if len(args) > 4 else m
    assert lines[1].split()[1:6] == [str(args[0]), str(args[1]), str(m), str(n), str(k)]
    for key in ("copy-in", "transpose:", "device:", "host:", "diff.cur:"):
        assert any(line.startswith(key) for line in lines), (key, out)
    gflops = float(re.search(r"device: \S+ ms (\S+) GFLOPS/s", out).group(1))
    diff = float(re.search(r"diff.cur: (\S+)", out).group(1))
    assert gflops > 0 and diff <= bar, out
