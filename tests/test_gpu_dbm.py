"""GPU parity of DBM task records on the stack kernels (opencl_libsmm_acc_process, param_format != 0).

A DBM task record is 6 ints (m, n, k, a, b, c) with 0-based element offsets (samples/smm/kernels/
multiply.cl:96-99: pnext = 6, pzero = 0, the offsets read from the record's fourth int). The same
entries as DBCSR triplets (3 ints, 1-based) must give bit-identical C through every kernel, and C must
meet the oracle bar: max normalised error <= 1e-13 (FP64) / 1e-5 (FP32).
"""
import numpy as np
import pytest

import libxstream_amd as lx
from libxstream_amd.workload import make_stack_workload
from tests.helpers import oracle_error, tol, uncut_blocks

pytestmark = pytest.mark.gpu

SHAPES = [(5, 5, 5), (13, 13, 13), (23, 23, 23), (4, 9, 22), (45, 45, 13), (64, 64, 64)]
DTYPES = [np.float64, np.float32]


@pytest.fixture(scope="module")
def stream():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    s = lx.Stream("dbm")
    yield s
    s.close()
    lx.finalize()


def dbm_records(w, stack=None):
    """DBCSR triplets (a, b, c), 1-based -> DBM task records (m, n, k, a, b, c), 0-based."""
    st = (w.stack if stack is None else stack).astype(np.int32)
    rec = np.empty((len(st), 6), dtype=np.int32)
    rec[:, 0], rec[:, 1], rec[:, 2] = w.m, w.n, w.k
    rec[:, 3:] = st - 1
    return rec


def run_both(stream, w, stack=None, event=None):
    """C after the triplet stack (param_format 0) and after the same entries as DBM records."""
    tid = lx.typeid(w.dtype)
    stack = w.stack if stack is None else stack
    bufs = []
    try:
        a = lx.DeviceBuffer.from_array(w.a)
        b = lx.DeviceBuffer.from_array(w.b_device_layout())
        st3 = lx.DeviceBuffer.from_array(np.ascontiguousarray(stack, dtype=np.int32))
        st6 = lx.DeviceBuffer.from_array(dbm_records(w, stack))
        bufs += [a, b, st3, st6]
        outs = []
        for st, fmt in ((st3, 0), (st6, lx.pack_format(w.m, w.n, w.k))):
            c = lx.DeviceBuffer.from_array(w.c)
            bufs.append(c)
            rc = lx.process_format(st, len(stack), tid, a, b, c, w.m, w.n, w.k, fmt, stream,
                                   event=(event if fmt else None))
            assert rc == 0, "opencl_libsmm_acc_process(param_format=%d) returned %d" % (fmt, rc)
            if fmt and event is not None:
                event.sync()
            stream.sync()
            out = np.empty_like(w.c)
            c.download(out)
            outs.append(out)
        return outs
    finally:
        for buf in bufs:
            buf.free()


def same_bits(w, out3, out6):
    """Both record formats run the same kernels: identical bits in every C block of a single entry
    (never cut); a run longer than 2 bs is cut and its pieces add atomically in either order."""
    short = uncut_blocks(w, limit=1)
    bound = (1e-12 if out3.dtype == np.float64 else 1e-5) * max(1.0, float(np.abs(out3).max()))
    return np.array_equal(out3[short], out6[short]) and np.allclose(out3[~short], out6[~short], rtol=0, atol=bound)


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
@pytest.mark.parametrize("shape", SHAPES, ids=["%dx%dx%d" % s for s in SHAPES])
def test_dbm_records(stream, shape, dtype):
    m, n, k = shape
    w = make_stack_workload(m, n, k, 2000, dtype=dtype, seed=hash((shape, "dbm")) % 2**31, c_init="uniform")
    out3, out6 = run_both(stream, w)
    assert same_bits(w, out3, out6), "DBM records differ from triplets"
    err = oracle_error(w, out6)
    assert err[0] <= tol(dtype), "max normalised error %g" % err[0]


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
def test_dbm_every_variant(stream, dtype):
    """Each compiled kernel kind (vector FMA with LDS, wavefront-per-run, matrix cores) reads DBM records."""
    tid = lx.typeid(dtype)
    checked = 0
    for shape in [(5, 5, 5), (7, 11, 3), (23, 23, 23)]:
        m, n, k = shape
        w = make_stack_workload(m, n, k, 600, dtype=dtype, seed=7 + m, c_init="uniform")
        saved, _ = lx.config_get(tid, m, n, k)
        try:
            for variant, tm, tn, nt, vkt in lx.kernel_variants(tid):
                for flush in (1, 2):
                    cfg = lx.SmmConfig(variant=variant, tm=tm, tn=tn, nt=nt, bs=16, kt=vkt if variant else 0,
                                       flush=flush)
                    if not lx.config_valid(tid, m, n, k, cfg):
                        continue
                    lx.config_set(tid, m, n, k, cfg)
                    out3, out6 = run_both(stream, w)
                    assert same_bits(w, out3, out6), "%s %r" % (shape, cfg)
                    checked += 1
            # the kernel compiled for this shape (5^3: wavefront-per-run; others: the tuned one)
            lx.config_set(tid, m, n, k, saved)
            out3, out6 = run_both(stream, w)
            assert same_bits(w, out3, out6)
            assert oracle_error(w, out6)[0] <= tol(dtype)
        finally:
            lx.config_set(tid, m, n, k, saved)
    assert checked >= 20


def test_dbm_unsorted_atomic(stream):
    """Records not grouped by C: the atomic flush still adds every entry once."""
    m = n = k = 13
    w = make_stack_workload(m, n, k, 1500, dtype=np.float64, seed=11, c_init="uniform")
    rng = np.random.default_rng(3)
    stack = w.stack[rng.permutation(len(w.stack))]
    out3, out6 = run_both(stream, w, stack=stack)
    assert oracle_error(w, out6, stack=stack)[0] <= 1e-13
    assert oracle_error(w, out3, stack=stack)[0] <= 1e-13


def test_dbm_event(stream):
    """The event passed to opencl_libsmm_acc_process completes with the launch."""
    m = n = k = 23
    w = make_stack_workload(m, n, k, 1000, dtype=np.float64, seed=5, c_init="uniform")
    ev = lx.Event()
    try:
        out3, out6 = run_both(stream, w, event=ev)
    finally:
        ev.close()
    assert same_bits(w, out3, out6)


def test_dbm_fallback_codes(stream):
    """Same -1 / no-op contract as libsmm_acc_process."""
    a = lx.DeviceBuffer(8)
    try:
        fmt = lx.pack_format(23, 23, 23)
        assert lx.process_format(a, 0, lx.REAL_8, a, a, a, 23, 23, 23, fmt, stream) == 0
        assert lx.process_format(a, 10, lx.REAL_8, a, a, a, 23, 23, 23, fmt, stream, def_mnk=0) == -1
        assert lx.process_format(a, 10, 5, a, a, a, 23, 23, 23, fmt, stream) == -1
        assert lx.process_format(a, 10, lx.REAL_8, a, a, a, 90, 90, 23, fmt, stream) == -1
    finally:
        a.free()
