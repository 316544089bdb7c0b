"""GPU parity of the Ozaki scheme II (CRT) GEMM (include/ozaki.h, kind 2 and the adaptive kind 3)
through the C ABI.

Bar: bit for bit equal to the oracle (oracle/ozaki_oracle.c: the exact integer product of the aligned
mantissas, rounded once), for every trans combination, ragged and tile-aligned shapes, FP32, trimmed
mantissas, K passes (explicit, and the ones a reduced prime count forces), small work items (many
residue partials per prime), host operands and zero/subnormal/non-finite inputs. The device gets
there through int8 residues, one int8 GEMM per modulus, Garner's mixed-radix digits and a 128-bit
Horner evaluation, none of which the oracle uses.
Accuracy against the double-double GEMM is stated beside the prime count: 16 moduli (125 bits)
with per-row/column exponents give a normwise error near 1e-16.
"""
import numpy as np
import pytest

import libxstream_amd as lx
import oracle

pytestmark = pytest.mark.gpu

# the library's moduli in use order (LXS_OZ_MODULI): pairwise coprime, at most 256
MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191]


def kmax(nprimes, bits):
    """longest K pass nprimes moduli keep exact: 2 K (2^bits - 1)^2 < M"""
    M = 1
    for m in MODULI[:nprimes]:
        M *= m
    return (M - 1) // (2 * (2 ** bits - 1) ** 2)


@pytest.fixture(scope="module")
def stream():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    s = lx.Stream("ozaki2")
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


def _expect(ta, tb, m, n, k, alpha, args, beta, trim=0, kpass=0):
    a, lda, b, ldb, c = args
    return oracle.ozaki2_gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, m, trim=trim, kpass=kpass)


def _mixed(rng, m, k):
    return rng.uniform(-1, 1, (m, k)) * np.exp2(rng.integers(-30, 31, (m, k)))


SHAPES = [(1, 1, 1), (17, 33, 65), (128, 128, 64), (200, 130, 300), (257, 129, 511)]


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("shape", SHAPES)
def test_crt_bitexact_trans_shapes(stream, ta, tb, shape):
    m, n, k = shape
    rng = np.random.default_rng(m * 5 + n * 7 + k)
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(kind=2)
    try:
        out, args = _run(oz, stream, ta, tb, m, n, k, 1.5, A, B, -0.5, C0)
    finally:
        oz.close()
    assert np.array_equal(out, _expect(ta, tb, m, n, k, 1.5, args, -0.5))


@pytest.mark.parametrize("trim", [0, 3])
def test_crt_bitexact_fp32(stream, trim):
    rng = np.random.default_rng(21)
    m, n, k = 150, 97, 333
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(use_double=False, kind=2, trim=trim)
    try:
        out, args = _run(oz, stream, "N", "T", m, n, k, 0.75, A, B, 2.0, C0, dtype=np.float32)
    finally:
        oz.close()
    assert out.dtype == np.float32
    assert np.array_equal(out, _expect("N", "T", m, n, k, 0.75, args, 2.0, trim=2 * trim))


@pytest.mark.parametrize("trim", [1, 5, 20])
def test_crt_bitexact_trim(stream, trim):
    """kind 2 drops 2 trim mantissa bits (samples/ozaki/ozaki_opencl.c:428-431), fewer primes follow"""
    rng = np.random.default_rng(22 + trim)
    m, n, k = 64, 300, 190
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(kind=2, trim=trim)
    try:
        out, args = _run(oz, stream, "T", "N", m, n, k, 1.0, A, B, 0.0, C0)
    finally:
        oz.close()
    assert np.array_equal(out, _expect("T", "N", m, n, k, 1.0, args, 0.0, trim=2 * trim))


def test_crt_k_passes_and_small_work_items(stream, monkeypatch):
    """explicit K passes, and 3 k-blocks per work item: many residue bytes per prime summed mod m"""
    monkeypatch.setenv("LIBXSTREAM_OZAKI_FLUSH", "3")
    rng = np.random.default_rng(23)
    m, n, k = 260, 140, 1000
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(kind=2, maxk=300)
    try:
        out, args = _run(oz, stream, "N", "N", m, n, k, -1.0, A, B, 0.5, C0)
    finally:
        oz.close()
    assert np.array_equal(out, _expect("N", "N", m, n, k, -1.0, args, 0.5, kpass=300))


def test_crt_reduced_primes_split_k(stream):
    """15 moduli keep only K <= kmax(15) exact: the library splits K into such passes"""
    kp = kmax(15, 53)
    assert 1000 < kp < 2000
    rng = np.random.default_rng(24)
    m, n, k = 100, 90, 2 * kp + 77
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(kind=2, nprimes=15)
    try:
        out, args = _run(oz, stream, "N", "N", m, n, k, 1.0, A, B, 1.0, C0)
    finally:
        oz.close()
    assert np.array_equal(out, _expect("N", "N", m, n, k, 1.0, args, 1.0, kpass=kp))


def test_crt_too_few_primes_fails(stream):
    rng = np.random.default_rng(25)
    oz = lx.Ozaki(kind=2, nprimes=8)  # 64 bits cannot hold a product of two 53-bit mantissas
    try:
        with pytest.raises(RuntimeError):
            _run(oz, stream, "N", "N", 8, 8, 8, 1.0, _mixed(rng, 8, 8), _mixed(rng, 8, 8), 0.0, np.zeros((8, 8)))
    finally:
        oz.close()


def test_crt_host_operands(stream):
    rng = np.random.default_rng(26)
    m, n, k = 70, 45, 129
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(kind=2)
    try:
        out, args = _run(oz, stream, "N", "N", m, n, k, 2.0, A, B, -1.0, C0, dev=False)
    finally:
        oz.close()
    assert np.array_equal(out, _expect("N", "N", m, n, k, 2.0, args, -1.0))


def test_crt_special_values_and_degenerate(stream):
    rng = np.random.default_rng(27)
    m, n, k = 33, 20, 70
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    A[0, :5] = [0.0, -0.0, 5e-324, np.inf, np.nan]
    A[3, :] = 0.0
    B[:, 2] = 1e-310
    B[7, :] = -np.inf
    oz = lx.Ozaki(kind=2)
    try:
        out, args = _run(oz, stream, "N", "N", m, n, k, 1.0, A, B, 0.25, C0)
        assert np.array_equal(out, _expect("N", "N", m, n, k, 1.0, args, 0.25))
        for (mm, nn, kk, al, be) in ((m, n, 0, 1.0, 0.5), (m, n, k, 0.0, 2.0), (m, n, 0, 1.0, 0.0)):
            out, args = _run(oz, stream, "N", "N", mm, nn, kk, al, A[:, :kk], B[:kk, :], be, C0)
            assert np.array_equal(out, _expect("N", "N", mm, nn, kk, al, args, be))
    finally:
        oz.close()


@pytest.mark.parametrize("k,scheme", [(8, 1), (512, 2)])
def test_adaptive_kind3(stream, k, scheme):
    """kind 3 takes scheme I (8 slices, 64 pairs) at tiny K and scheme II otherwise"""
    rng = np.random.default_rng(28 + k)
    m, n = 96, 80
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(kind=3)
    try:
        out, (a, lda, b, ldb, c) = _run(oz, stream, "N", "N", m, n, k, 1.0, A, B, 0.5, C0)
    finally:
        oz.close()
    if 1 == scheme:
        ref = oracle.ozaki_gemm("N", "N", m, n, k, 1.0, a, lda, b, ldb, 0.5, c, m, 8)
    else:
        ref = oracle.ozaki2_gemm("N", "N", m, n, k, 1.0, a, lda, b, ldb, 0.5, c, m)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("e", [0, 10, 30])
def test_crt_accuracy_rowcol_exponents(stream, e):
    rng = np.random.default_rng(29 + e)
    m = n = k = 256
    A = rng.uniform(1, 2, (m, k)) * rng.choice([-1, 1], (m, k)) * np.exp2(rng.integers(-e, e + 1, (m, 1)))
    B = rng.uniform(1, 2, (k, n)) * rng.choice([-1, 1], (k, n)) * np.exp2(rng.integers(-e, e + 1, (1, n)))
    oz = lx.Ozaki(kind=2)
    try:
        out, (a, lda, b, ldb, _) = _run(oz, stream, "N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n)))
    finally:
        oz.close()
    hi, lo = oracle.gemm_dd("N", "N", m, n, k, 1.0, a, lda, b, ldb, 0.0, None, m)
    err = np.linalg.norm(out - (hi + lo)) / np.linalg.norm(hi + lo)
    print("scheme II, 16 moduli, e=%d: normwise error %.3e" % (e, err))
    assert err <= 5e-16


def test_crt_complex_3m(stream):
    rng = np.random.default_rng(30)
    m, n, k = 60, 50, 90
    A = rng.uniform(-1, 1, (m, k)) + 1j * rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n)) + 1j * rng.uniform(-1, 1, (k, n))
    C = rng.uniform(-1, 1, (m, n)) + 1j * rng.uniform(-1, 1, (m, n))
    alpha, beta = 1.0 - 0.5j, 0.25 + 1j
    inter = lambda X: np.asfortranarray(X).ravel(order="F").view(np.float64).copy()
    da, db, dc = (lx.DeviceBuffer.from_array(inter(x)) for x in (A, B, C))
    oz = lx.Ozaki(kind=2)
    try:
        oz.gemm_complex("N", "N", m, n, k, alpha, da, m, db, k, beta, dc, m, stream=stream)
        stream.sync()
        out = np.empty(2 * m * n)
        dc.download(out)
    finally:
        oz.close()
        for x in (da, db, dc):
            x.free()
    got = out.view(np.complex128).reshape(n, m).T
    ref = alpha * (A @ B) + beta * C
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-14


@pytest.mark.parametrize("stages", [4, 5])
def test_crt_lds_stages(stream, monkeypatch, stages):
    """both pipeline depths of the int8 GEMM (4 or 5 LDS stages; 5 pads items to ten-stage trips)"""
    monkeypatch.setenv("LIBXSTREAM_OZAKI_STAGES", str(stages))
    rng = np.random.default_rng(50 + stages)
    m, n, k = 270, 300, 2100
    A, B, C0 = _mixed(rng, m, k), _mixed(rng, k, n), rng.uniform(-1, 1, (m, n))
    oz = lx.Ozaki(kind=2)
    try:
        out, args = _run(oz, stream, "T", "N", m, n, k, 1.0, A, B, -1.0, C0)
    finally:
        oz.close()
    assert np.array_equal(out, _expect("T", "N", m, n, k, 1.0, args, -1.0))
