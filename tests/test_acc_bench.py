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


def _reports(out):
    """(header fields, diff.cur) of every benchmark in a report"""
    lines = out.splitlines()
    heads = [i for i, line in enumerate(lines) if line.split(" ")[0].endswith(("acc_bench", "acc_bench_f32"))]
    res = []
    for i in heads:
        fields = [int(x) for x in lines[i].split()[1:]]
        diff = next((float(line.split()[1]) for line in lines[i:] if line.startswith("diff.cur:")), None)
        res.append((fields, diff))
    return res


@pytest.mark.gpu
def test_shape_file_and_diff_max(tmp_path):
    """a file of argument lists (blank lines skipped, "MxNxK" or "M N K"): one benchmark per line and diff.max
    over them (samples/smm/acc_bench.c:148, :442-462, :491-557)"""
    f = tmp_path / "shapes.txt"
    f.write_text("2 3000 23\n\n2 2000 4x9x22\n5x5x5 100\n")
    rc, out, err = _run(BIN, [f], env={"CHECK": "1"})
    assert rc == 0, out + err
    reps = _reports(out)
    assert [r[0][:5] for r in reps] == [[2, 3000, 23, 23, 23], [2, 2000, 4, 9, 22], [3, 30000, 5, 5, 5]], out
    assert reps[2][0][5] == 100  # nc from the line, na = nb = 10 nc
    diffs = [r[1] for r in reps]
    assert all(d is not None and d <= 1e-13 for d in diffs), out
    dmax = float(re.search(r"diff.max: (\S+)", out).group(1))
    assert dmax == pytest.approx(max(diffs), rel=1e-5, abs=1e-300)


@pytest.mark.gpu
def test_performance_floors():
    """CHECK_DEV / CHECK_HST / CHECK_H2D: the run fails when the device, host or copy-in rate is below the floor
    (samples/smm/acc_bench.c:463-477)"""
    rc, out, err = _run(BIN, [2, 3000, 23], env={"CHECK": "1", "CHECK_DEV": "1e9"})
    assert rc != 0 and "FAILED" in err, out + err
    rc, out, err = _run(BIN, [2, 3000, 23], env={"CHECK": "1", "CHECK_HST": "1e9"})
    assert rc != 0 and "FAILED" in err, out + err
    rc, out, err = _run(BIN, [2, 3000, 23], env={"CHECK": "1", "CHECK_H2D": "1e9"})
    assert rc != 0 and "FAILED" in err, out + err
    rc, out, err = _run(BIN, [2, 3000, 23], env={"CHECK": "1", "CHECK_DEV": "1", "CHECK_HST": "1e-3", "CHECK_H2D": "1e-3"})
    assert rc == 0, out + err


@pytest.mark.gpu
def test_stack_size_forms():
    """a k/m/g byte budget (the smallest stack whose pools and stack reach it), BATCHSIZE_SMM=<size>,<nrepeat>,
    a negative size (a random multiple of 100 up to its magnitude) and CHECK=0 (no host run)"""
    rc, out, err = _run(BIN, [2, "4m", 23], env={"CHECK": "1"})
    assert rc == 0, out + err
    (nr, s, m, n, k, nc, na, nb), diff = _reports(out)[0]

    def size(s):
        nc = max(s // 16, 1)
        return 8 * (m * k + k * n) * 10 * nc + 12 * s

    assert size(s) >= 4 << 20 > size(s - 1) and diff <= 1e-13, out
    rc, out, err = _run(BIN, [2, 500, 13], env={"CHECK": "1", "BATCHSIZE_SMM": "2000,4"})
    assert rc == 0 and _reports(out)[0][0][:2] == [4, 2000], out + err
    rc, out, err = _run(BIN, [2, -5000, 23], env={"CHECK": "1"})
    (nr, s, *_), diff = _reports(out)[0]
    assert rc == 0 and s % 100 == 0 and 100 <= s <= 5000 and nr >= 3 and diff <= 1e-13, out + err
    rc, out, err = _run(BIN, ["", 3000, 23], env={"CHECK": "0"})
    assert rc == 0 and "host:" not in out and _reports(out)[0][0][0] == 66, out + err
