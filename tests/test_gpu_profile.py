"""GPU checks of the launch and transfer profile (LIBXSTREAM_PROFILE / LIBXSTREAM_PROFILE_MEM).

- Each stack launch states the reference's work: 2 m n k S flops and S ts (mk + kn + 2mn) bytes
  (samples/smm/smm_kernel.c:504-506); each transpose 2 S m n ts bytes (samples/smm/smm_trans.c:208-212).
- Summaries and histograms: the median sample's rate, shortest <= median <= longest, one sample per launch.
- Copies and fills are recorded by kind with their sizes (src/libxstream_mem.c:805-893).
"""
import numpy as np
import pytest

import libxstream_amd as lx
from libxstream_amd.workload import make_stack_workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    s = lx.Stream("profile")
    yield s
    s.close()
    lx.finalize()


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("shape", [(23, 23, 23), (4, 9, 22), (45, 45, 13)], ids=lambda s: "%dx%dx%d" % s)
def test_launch_records_state_reference_work(stream, shape, dtype, capfd):
    m, n, k = shape
    w = make_stack_workload(m, n, k, 3000, dtype=dtype, seed=3)
    tid = lx.typeid(dtype)
    ts = np.dtype(dtype).itemsize
    a, b, c, st = (lx.DeviceBuffer.from_array(x) for x in (w.a, w.b, w.c, w.stack))
    trs = lx.DeviceBuffer.from_array(np.arange(w.nb, dtype=np.int32) * (k * n))
    try:
        lx.profile_reset()
        lx.profile_enable(True)
        for _ in range(7):
            assert lx.process(st, w.stack_size, tid, a, b, c, m, n, k, stream) == 0
        lx.transpose(trs, 0, w.nb, b, tid, k, n, stream)
        stream.sync()
        smm = lx.profile_get("smm")
        trans = lx.profile_get("trans")
        stats = lx.profile_stats("smm")
        lx.profile_print()
        lx.profile_enable(False)
    finally:
        for x in (a, b, c, st, trs):
            x.free()
    assert smm["count"] + stats["discarded"] == 7
    assert smm["flops"] == pytest.approx(smm["count"] * 2.0 * m * n * k * w.stack_size, rel=1e-12)
    assert smm["bytes"] == pytest.approx(smm["count"] * w.stack_size * ts * (m * k + k * n + 2 * m * n), rel=1e-12)
    assert trans["bytes"] == pytest.approx(trans["count"] * 2.0 * w.nb * k * n * ts, rel=1e-12)
    assert 0 < stats["min_ms"] <= stats["median_ms"] <= stats["max_ms"]
    assert stats["median_rate"] == pytest.approx(2.0 * m * n * k * w.stack_size / stats["median_ms"] * 1e-6, rel=1e-5)
    err = capfd.readouterr().err
    assert "PROF ACC/HIP: smm=" in err and "GFLOPS/s" in err


def test_transfer_records(stream, capfd):
    sizes = (1 << 20, 64 << 20)
    host = [lx.HostBuffer(np.uint8, s, stream) for s in sizes]
    dev = [lx.DeviceBuffer(s) for s in sizes]
    try:
        lx.profile_reset()
        lx.profile_mem_enable(True)
        for _ in range(3):
            for h, d, s in zip(host, dev, sizes):
                lx.check(lx.lib().libxstream_mem_copy_h2d(h.ptr, d.ptr, s, stream.handle))
                lx.check(lx.lib().libxstream_mem_copy_d2h(d.ptr, h.ptr, s, stream.handle))
        dev[1].zero(stream)
        stream.sync()
        h2d = lx.profile_get("H2D")
        d2h = lx.profile_get("D2H")
        zero = lx.profile_get("ZERO")
        stats = lx.profile_stats("H2D")
        lx.profile_print()
        lx.profile_mem_enable(False)
    finally:
        for x in host + dev:
            x.free()
    assert h2d["count"] + stats["discarded"] == 6 and d2h["count"] >= 1 and zero["count"] == 1
    assert h2d["bytes"] <= 3.0 * sum(sizes) and zero["bytes"] == sizes[1]
    assert stats["median_rate"] > 0.5  # GB/s from pinned memory
    err = capfd.readouterr().err
    assert "PROF ACC/HIP: H2D=" in err and "GB/s" in err
