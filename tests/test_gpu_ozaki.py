"""GPU parity of the Ozaki scheme I GEMM (include/ozaki.h) through the C ABI.

Bar: the HIP path equals the oracle's restatement (oracle/ozaki_oracle.c, same operation order) bit
for bit, for every trans combination, ragged and tile-aligned shapes, slice counts, trims, alpha/beta,
FP32, K passes, small work items (many int32 partials per level), host operands, and
zero/subnormal/non-finite inputs.
Accuracy against the double-double GEMM is stated with the slice count: per-row/column exponents
(mantissas uniform [1,2), exponents uniform in [-e, e]) give a normwise error <= 1e-14 with 7 slices
over the triangle and <= 1e-15 with 8. The native FP64 matrix-core GEMM (the comparator) and the
complex 3M path are checked against the double-double GEMM.
"""

import numpy as np
import pytest

import libxstream_amd as lx
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    s = lx.Stream("ozaki")
    yield s
    s.close()
    lx.finalize()


def _store(X, trans):
    S = X.T if trans == "T" else X
    return np.asfortranarray(S).ravel(order="F"), max(1, S.shape[0])


def _run(oz, stream, ta, tb, m, n, k, alpha, A, B, beta, C0, dtype=np.float64, dev=True):
    a, lda = _store(A, ta)
    b, ldb = _store(B, tb)
    a, b = a.astype(dtype), b.astype(dtype)
    c = np.asfortranarray(C0).ravel(order="F").astype(dtype)
    if dev:
        da, db, dc = (lx.DeviceBuffer.from_array(x) for x in (a, b, c))
        oz.gemm(ta, tb, m, n, k, alpha, da, lda, db, ldb, beta, dc, m, stream=stream)
        stream.sync()
        out = np.empty_like(c)
        dc.download(out)
        for x in (da, db, dc):
            x.free()
    else:
        out = c.copy()
        oz.gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, out, m, stream=stream, dev=False)
        stream.sync()
    return out, (a, lda, b, ldb, c)


def _expect(ta, tb, m, n, k, alpha, args, beta, s, trim=0, maxk=0):
    a, lda, b, ldb, c = args
    return oracle.ozaki_gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, m, s, trim=trim, maxk=maxk)


def _mixed(rng, m, k):
    return rng.uniform(-1, 1, (m, k)) * np.exp2(rng.integers(-30, 31, (m, k)))


SHAPES = [(1, 1, 1), (17, 33, 65), (128, 128, 64), (200, 130, 300), (257, 129, 511)]


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("shape", SHAPES)
def test_bitexact_trans_shapes(stream, ta, tb, shape):
    m, n, k = shape
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(nslices=7, trim=6)
    try:
        out, args = _run(oz, stream, ta, tb, m, n, k, 1.5, A, B, -0.5, C0)
    finally:
        oz.close()
    assert np.array_equal(out, _expect(ta, tb, m, n, k, 1.5, args, -0.5, 7, trim=6))


@pytest.mark.parametrize("s,trim", [(1, 0), (2, 0), (3, 2), (5, 0), (7, 0), (8, 0), (8, 7), (8, 3)])
def test_bitexact_slices_trims(stream, s, trim):
    m, n, k = 150, 140, 200
    rng = np.random.default_rng(s * 10 + trim)
    A, B = _mixed(rng, m, k), _mixed(rng, k, n)
    oz = lx.Ozaki(nslices=s, trim=trim)
    try:
        out, args = _run(oz, stream, "N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n)))
    finally:
        oz.close()
    assert np.array_equal(out, _expect("N", "N", m, n, k, 1.0, args, 0.0, s, trim=trim))


def test_bitexact_fp32(stream):
    m, n, k = 100, 90, 130
    rng = np.random.default_rng(32)
    A, B, C0 = rng.uniform(-4, 4, (m, k)), rng.uniform(-4, 4, (k, n)), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(use_double=False)
    try:
        assert oz.ctx.nslices == 4
        out, args = _run(oz, stream, "N", "T", m, n, k, 1.0, A, B, 1.0, C0, dtype=np.float32)
    finally:
        oz.close()
    exp = _expect("N", "T", m, n, k, 1.0, args, 1.0, 4)
    assert out.dtype == np.float32 and np.array_equal(out, exp)


def test_bitexact_k_passes_and_small_work_items(stream, monkeypatch):
    """K in passes of 300, work items of 3 k-blocks: level sums stay exact, so the result is the same"""
    m, n, k = 130, 70, 700
    rng = np.random.default_rng(77)
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    monkeypatch.setenv("LIBXSTREAM_OZAKI_FLUSH", "3")
    oz = lx.Ozaki(nslices=6, trim=2, maxk=300)
    try:
        assert oz.ctx.flush == 3 and oz.ctx.maxk == 300
        out, args = _run(oz, stream, "T", "N", m, n, k, -1.0, A, B, 2.0, C0)
    finally:
        oz.close()
    assert np.array_equal(out, _expect("T", "N", m, n, k, -1.0, args, 2.0, 6, trim=2, maxk=300))


def test_bitexact_host_operands(stream):
    m, n, k = 96, 200, 77
    rng = np.random.default_rng(5)
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki()
    try:
        out, args = _run(oz, stream, "N", "N", m, n, k, 1.0, A, B, 1.0, C0, dev=False)
    finally:
        oz.close()
    assert np.array_equal(out, _expect("N", "N", m, n, k, 1.0, args, 1.0, 8))


def test_special_values_and_degenerate(stream):
    m, n, k = 64, 64, 64
    rng = np.random.default_rng(6)
    A, B = rng.uniform(-1, 1, (m, k)), rng.uniform(-1, 1, (k, n))
    A[3, :] = 0.0                  # zero row: no exponent
    A[5, 7] = 4e-320               # subnormal counts as zero
    B[:, 9] = 0.0                  # zero column
    A[10, 11], B[12, 13] = np.inf, np.nan  # non-finite count as zero (ozaki_common.cl:784-805)
    C0 = rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(nslices=7, trim=6)
    try:
        out, args = _run(oz, stream, "N", "N", m, n, k, 1.0, A, B, 1.0, C0)
        assert np.array_equal(out, _expect("N", "N", m, n, k, 1.0, args, 1.0, 7, trim=6))
        out, _ = _run(oz, stream, "N", "N", m, n, k, 0.0, A, B, 3.0, C0)   # alpha = 0: C = beta C
        assert np.array_equal(out, 3.0 * np.asfortranarray(C0).ravel(order="F"))
        out, _ = _run(oz, stream, "N", "N", m, n, 0, 1.0, A[:, :0], B[:0, :], 0.0, C0)  # K = 0
        assert not out.any()
    finally:
        oz.close()


@pytest.mark.parametrize("e", [0, 10, 30])
def test_accuracy_rowcol_exponents(stream, e):
    m = n = k = 512
    rng = np.random.default_rng(1000 + e)
    A = rng.uniform(1, 2, (m, k)) * rng.choice([-1.0, 1.0], (m, k)) * np.exp2(rng.integers(-e, e + 1, (m, 1)))
    B = rng.uniform(1, 2, (k, n)) * rng.choice([-1.0, 1.0], (k, n)) * np.exp2(rng.integers(-e, e + 1, (1, n)))
    a, b = np.asfortranarray(A).ravel(order="F"), np.asfortranarray(B).ravel(order="F")
    hi, lo = oracle.gemm_dd("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, None, m)
    for s, trim, bound in ((7, 6, 1e-14), (8, 7, 1e-15), (8, 0, 1e-15)):
        oz = lx.Ozaki(nslices=s, trim=trim)
        try:
            out, _ = _run(oz, stream, "N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n)))
        finally:
            oz.close()
        err = oracle.error(out, hi, lo)[1]
        print("ozaki e=%d s=%d trim=%d normwise error %.2e" % (e, s, trim, err))
        assert err <= bound


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_native_dgemm(stream, ta, tb):
    m, n, k = 300, 260, 333
    rng = np.random.default_rng(3)
    A, B, C0 = rng.uniform(-1, 1, (m, k)), rng.uniform(-1, 1, (k, n)), rng.uniform(-1, 1, (m, n))
    a, lda = _store(A, ta)
    b, ldb = _store(B, tb)
    c = np.asfortranarray(C0).ravel(order="F")
    da, db, dc = (lx.DeviceBuffer.from_array(x) for x in (a, b, c))
    lx.dgemm_native(ta, tb, m, n, k, 2.0, da, lda, db, ldb, -1.0, dc, m, stream=stream)
    stream.sync()
    out = np.empty_like(c)
    dc.download(out)
    for x in (da, db, dc):
        x.free()
    hi, lo = oracle.gemm_dd(ta, tb, m, n, k, 2.0, a, lda, b, ldb, -1.0, c, m)
    assert oracle.error(out, hi, lo)[0] <= 1e-14


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "T")])
def test_complex_3m(stream, ta, tb):
    m, n, k = 70, 90, 110
    rng = np.random.default_rng(8)
    A = rng.uniform(-1, 1, (m, k)) + 1j * rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n)) + 1j * rng.uniform(-1, 1, (k, n))
    C0 = rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n))
    SA = A.conj().T if ta == "C" else A
    SB = B.T if tb == "T" else B
    alpha, beta = 1.5 - 0.5j, 0.25 + 1j
    da = lx.DeviceBuffer.from_array(np.asfortranarray(SA).ravel(order="F").view(np.float64))
    db = lx.DeviceBuffer.from_array(np.asfortranarray(SB).ravel(order="F").view(np.float64))
    dc = lx.DeviceBuffer.from_array(np.asfortranarray(C0).ravel(order="F").view(np.float64))
    oz = lx.Ozaki()
    try:
        oz.gemm_complex(ta, tb, m, n, k, alpha, da, SA.shape[0], db, SB.shape[0], beta, dc, m, stream=stream)
        stream.sync()
    finally:
        oz.close()
    out = np.empty(2 * m * n)
    dc.download(out)
    for x in (da, db, dc):
        x.free()
    got = out.view(np.complex128).reshape(n, m).T
    ref = alpha * (A @ B) + beta * C0  # numpy complex128 product: rounding ~1e-16, far below the bound
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-13


@pytest.mark.parametrize("stages", [4, 5])
def test_bitexact_lds_stages(stream, monkeypatch, stages):
    """both pipeline depths of the int8 GEMM (4 or 5 LDS stages; 5 pads items to ten-stage trips)"""
    monkeypatch.setenv("LIBXSTREAM_OZAKI_STAGES", str(stages))
    rng = np.random.default_rng(40 + stages)
    m, n, k = 300, 260, 1500
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(nslices=6, trim=2)
    try:
        out, args = _run(oz, stream, "N", "T", m, n, k, 1.0, A, B, 0.5, C0)
    finally:
        oz.close()
    assert np.array_equal(out, _expect("N", "T", m, n, k, 1.0, args, 0.5, 6, trim=2))
