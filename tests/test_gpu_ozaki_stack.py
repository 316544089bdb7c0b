"""GPU parity of the Ozaki scheme I stack path (include/ozaki.h: ozaki_stack) through the C ABI.

Bar: bit for bit equal to the oracle (oracle.ozaki_stack: the scheme I GEMM oracle applied entry by
entry in stack order) for square, odd and non-square blocks up to 64 x 64 x 64, slice counts and trims,
FP32, a stack that is not grouped by C (one work-group fallback), offsets that are not block starts
(skipped) and empty stacks. At BASELINE's size (64^3, 100k entries) the error against the
double-double stack oracle is stated beside the slice count.
"""
import numpy as np
import pytest

import libxstream_amd as lx
import oracle
from libxstream_amd.workload import make_stack_workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    s = lx.Stream("ozaki_stack")
    yield s
    s.close()
    lx.finalize()


def _run(stream, w, s, trim=0, use_double=True, stack=None):
    st = w.stack if stack is None else stack
    bl = w.b_device_layout()
    da, db, dc = (lx.DeviceBuffer.from_array(np.ascontiguousarray(x).reshape(-1)) for x in (w.a, bl, w.c))
    ds = lx.DeviceBuffer.from_array(np.ascontiguousarray(st, dtype=np.int32).reshape(-1))
    oz = lx.Ozaki(use_double=use_double, nslices=s, trim=trim)
    try:
        oz.stack(ds, len(st), w.m, w.n, w.k, da, w.a.shape[0], db, bl.shape[0], dc, stream)
        stream.sync()
        out = np.empty(w.c.size, dtype=w.c.dtype)
        dc.download(out)
    finally:
        oz.close()
        for x in (da, db, dc, ds):
            x.free()
    return out, bl


@pytest.mark.parametrize("shape,s,trim", [((64, 64, 64), 7, 6), ((23, 23, 23), 8, 0), ((4, 9, 22), 5, 1),
                                           ((45, 45, 13), 8, 7), ((1, 1, 1), 3, 0), ((64, 17, 33), 6, 2)])
def test_stack_bitexact(stream, shape, s, trim):
    m, n, k = shape
    w = make_stack_workload(m, n, k, 400, dtype=np.float64, seed=sum(shape) + s, c_init="random")
    out, bl = _run(stream, w, s, trim)
    assert np.array_equal(out, oracle.ozaki_stack(m, n, k, w.stack, w.a, bl, w.c, s, trim=trim))


def test_stack_bitexact_fp32(stream):
    m, n, k = 13, 13, 13
    w = make_stack_workload(m, n, k, 500, dtype=np.float32, seed=7, c_init="random")
    out, bl = _run(stream, w, 4, use_double=False)
    ref = oracle.ozaki_stack(m, n, k, w.stack, w.a, bl, w.c, 4, dtype=np.float32)
    assert out.dtype == np.float32 and np.array_equal(out, ref)


def test_stack_unsorted_and_bad_offsets(stream):
    m, n, k = 16, 24, 40
    w = make_stack_workload(m, n, k, 300, dtype=np.float64, seed=9, c_init="random")
    rng = np.random.default_rng(9)
    st = w.stack[rng.permutation(len(w.stack))].copy()  # C blocks revisited in no order
    st[5, 0] += 1      # not a block start: skipped
    st[17, 1] = (w.b.shape[0] + 3) * k * n + 1  # beyond the B pool: skipped
    out, bl = _run(stream, w, 6, stack=st)
    keep = np.ones(len(st), dtype=bool)
    keep[[5, 17]] = False
    assert np.array_equal(out, oracle.ozaki_stack(m, n, k, st[keep], w.a, bl, w.c, 6))


def test_stack_empty_and_single(stream):
    w = make_stack_workload(64, 64, 64, 1, dtype=np.float64, seed=3, c_init="random")
    out, bl = _run(stream, w, 7, trim=6)
    assert np.array_equal(out, oracle.ozaki_stack(64, 64, 64, w.stack, w.a, bl, w.c, 7, trim=6))
    out, _ = _run(stream, w, 7, stack=np.zeros((0, 3), dtype=np.int32))
    assert np.array_equal(out, w.c.reshape(-1))


@pytest.mark.parametrize("e", [0, 10, 30])
def test_stack_64cubed_100k_error(stream, e):
    """the batched 64^3 stack of 100k entries over pools of 2000 A, 2000 B and 1000 C blocks, inputs of the
    Ozaki workload (A blocks: mantissas +-[1, 2) times 2^x per row, B blocks per column, x uniform in
    [-e, e]): normwise error against the double-double stack oracle over the C blocks of the first 40 runs
    (a sorted stack: their entries are a prefix of it), <= 1e-14 with 7 slices and <= 1e-15 with 8"""
    m = n = k = 64
    w = make_stack_workload(m, n, k, 100000, nc=1000, na=2000, nb=2000, dtype=np.float64, seed=25071975)
    rng = np.random.default_rng(20260701 + e)
    sgn = lambda shape: np.where(rng.integers(0, 2, shape, dtype=np.int8) != 0, 1.0, -1.0)
    A = rng.uniform(1, 2, (2000, k, m)) * sgn((2000, k, m)) * np.exp2(rng.integers(-e, e + 1, (2000, 1, m)))
    Bh = rng.uniform(1, 2, (2000, n, k)) * sgn((2000, n, k)) * np.exp2(rng.integers(-e, e + 1, (2000, n, 1)))
    w.a = np.ascontiguousarray(A.reshape(2000, -1))  # column-major m x k blocks: row exponent on A(i, :)
    w.b = np.ascontiguousarray(Bh.reshape(2000, -1))  # column-major k x n blocks: column exponent on B(:, j)
    cblk = (w.stack[:, 2] - 1) // (m * n)
    head = int(np.searchsorted(cblk, 40))
    errs = {}
    for s, trim in ((8, 7), (7, 6)):
        out, bl = _run(stream, w, s, trim)
        if not errs:
            hi, lo = oracle.stack_dd(m, n, k, w.stack[:head], w.a, bl, w.c, b_rowmajor=True)
            hi, lo = hi.reshape(-1)[:40 * m * n], lo.reshape(-1)[:40 * m * n]
        d = (out[:40 * m * n] - hi) - lo
        errs[(s, trim)] = float(np.sqrt((d * d).sum() / (hi * hi).sum()))
    print("ozaki_stack 64^3 S=100k e=%d normwise error:" % e, {"s=%d trim=%d" % key: "%.2e" % v for key, v in errs.items()})
    assert errs[(8, 7)] <= 1e-15 and errs[(7, 6)] <= 1e-14
