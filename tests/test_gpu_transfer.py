"""GPU: per-call parameter-stack copies through pinned staging, overlapped with the stack kernels
(libxstream_amd.transfer.StackPipeline over libxstream_mem_copy_h2d, events and stream waits).

Every pipeline order must give the oracle's C for the concatenation of the stacks it was fed
(libsmm_acc_process bar: max normalised error <= 1e-13 FP64 / 1e-5 FP32).
"""
import numpy as np
import pytest

import libxstream_amd as lx
from libxstream_amd.transfer import StackPipeline
from libxstream_amd.workload import make_stack_workload
from tests.helpers import oracle_error, tol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def streams():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    run, up = lx.Stream("run"), lx.Stream("upload")
    yield run, up
    up.close()
    run.close()
    lx.finalize()


def _case(dtype, sizes, seed):
    m, n, k = 23, 23, 23
    w = make_stack_workload(m, n, k, 3000, dtype=dtype, seed=seed, c_init="uniform")
    stacks = [make_stack_workload(m, n, k, s, nc=w.nc, na=w.na, nb=w.nb, dtype=dtype, seed=seed + 1 + i).stack
              for i, s in enumerate(sizes)]
    return w, stacks


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("nslots,two_streams", [(1, False), (2, True), (3, True)], ids=["serial", "ring2", "ring3"])
def test_pipeline_parity(streams, dtype, nslots, two_streams):
    run, up = streams
    w, stacks = _case(dtype, [3000, 17, 2999, 1, 1500, 3000, 640], seed=101 + nslots)
    tid = lx.typeid(dtype)
    a = lx.DeviceBuffer.from_array(w.a)
    b = lx.DeviceBuffer.from_array(w.b_device_layout())
    c = lx.DeviceBuffer.from_array(w.c)
    pipe = StackPipeline(3000, run, up if two_streams else None, nslots=nslots)
    try:
        def launch(dev, entries):
            rc = lx.process(dev, entries, tid, a, b, c, w.m, w.n, w.k, run)
            assert rc == 0

        for st in stacks:
            pipe.submit(st, launch)
        pipe.sync()
        out = np.empty_like(w.c)
        c.download(out)
    finally:
        pipe.close()
        for x in (a, b, c):
            x.free()
    assert oracle_error(w, out, stack=np.concatenate(stacks))[0] <= tol(dtype)


def test_pipeline_rejects_oversize(streams):
    run, up = streams
    pipe = StackPipeline(10, run, up)
    try:
        with pytest.raises(ValueError):
            pipe.submit(np.ones((11, 3), dtype=np.int32), lambda dev, n: None)
    finally:
        pipe.close()
