"""Pins of the oracle's Ozaki scheme II (CRT) restatement (oracle/ozaki_oracle.c), CPU only.

The oracle forms the reference's integers (samples/ozaki/kernels/ozaki2_int8.cl:1808-1860: biased
exponents, mantissa with its implicit bit shifted by the exponent difference plus the trimmed bits)
and sums their products exactly, without residues. What fixes it:
  - exact products: integer matrices below 2^53 are aligned without loss, so the result must equal
    the int64 product bit for bit (trans combinations, alpha/beta, fp32);
  - an independent restatement with Python's exact rationals and integers in the other exponent
    convention (|v| < 2^E, a = floor(|v| 2^(mb - t - E))), rounded once by int -> float, with
    beta applied through an exactly rounded fused multiply-add (Fraction arithmetic);
  - the rigorous truncation bound against the double-double GEMM;
  - planted mistakes (rounded mantissas, dropped sign, scale off by one, FP64 summation), each of
    which one of the pins must reject.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle

U = 2.0 ** -53


def _fortran(x):
    return np.asfortranarray(x).ravel(order="F")


def _op_store(X, trans):
    S = X.T if trans == "T" else X
    return _fortran(S), S.shape[0]


def _rowcol(rng, m, k, e, rows, dtype=np.float64):
    mant = rng.uniform(1, 2, size=(m, k)) * rng.choice([-1.0, 1.0], size=(m, k))
    ex = rng.integers(-e, e + 1, size=(m, k))
    ex = ex + rng.integers(-e, e + 1, size=(m, 1) if rows else (1, k))
    return (mant * np.exp2(ex)).astype(dtype)


# ---- exact products --------------------------------------------------------------------------

@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_integer_products_exact(ta, tb):
    rng = np.random.default_rng(11)
    m, n, k = 7, 5, 33
    A = rng.integers(-2 ** 20, 2 ** 20, size=(m, k)).astype(np.float64)
    B = rng.integers(-2 ** 20, 2 ** 20, size=(k, n)).astype(np.float64)
    A[2, :] = 0  # an empty row
    B[:, 1] = 0  # an empty column
    a, lda = _op_store(A, ta)
    b, ldb = _op_store(B, tb)
    exact = A.astype(np.int64) @ B.astype(np.int64)  # |sum| < 2^46: exact in FP64
    out = oracle.ozaki2_gemm(ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, np.zeros(m * n), m)
    assert np.array_equal(out, _fortran(exact.astype(np.float64)))
    C = rng.integers(-1000, 1000, size=(m, n)).astype(np.float64)
    out = oracle.ozaki2_gemm(ta, tb, m, n, k, -2.0, a, lda, b, ldb, 0.5, _fortran(C), m)
    assert np.array_equal(out, _fortran(-2.0 * exact.astype(np.float64) + 0.5 * C))


def test_integer_products_exact_fp32_and_passes():
    rng = np.random.default_rng(12)
    m, n, k = 6, 9, 50
    A = rng.integers(-2 ** 10, 2 ** 10, size=(m, k)).astype(np.float32)
    B = rng.integers(-2 ** 10, 2 ** 10, size=(k, n)).astype(np.float32)
    exact = A.astype(np.int64) @ B.astype(np.int64)  # |sum| < 2^24: exact in FP32
    for kpass in (0, 7, 50):
        out = oracle.ozaki2_gemm("N", "N", m, n, k, 1.0, _fortran(A), m, _fortran(B), k, 0.0,
                                 np.zeros(m * n, np.float32), m, kpass=kpass)
        assert out.dtype == np.float32 and np.array_equal(out, _fortran(exact.astype(np.float32)))


# ---- exact rationals ----------------------------------------------------------------------------

def _expo(v):
    """E with 2^(E-1) <= |v| < 2^E for a normal finite v, None otherwise"""
    if v == 0 or not math.isfinite(v) or abs(v) < (2.2250738585072014e-308 if isinstance(v, float) else 0):
        return None
    return math.frexp(v)[1]


def _restated(A, B, alpha, beta, C, trim, mb, kpass, to_dtype):
    """the same contract with exact rationals: per K pass, a = floor(|v| 2^(mb - t - E)) with the sign,
    x = sum of products (Python int), rounded once, scaled by 2^(Ea + Eb + 2t - 2 mb), then
    fma(alpha, ., beta C) rounded once (Fraction), per element"""
    m, k = A.shape
    n = B.shape[1]
    tiny = 2.2250738585072014e-308 if mb == 53 else float(np.finfo(np.float32).tiny)
    out = C.astype(np.float64).copy()
    kpass = kpass or k
    for k0 in range(0, k, kpass):
        kk = min(kpass, k - k0)
        beta_eff = beta if k0 == 0 else 1.0

        def ints(rowvals):
            vals = [float(v) for v in rowvals]
            ok = [v != 0 and math.isfinite(v) and abs(v) >= tiny for v in vals]
            es = [math.frexp(v)[1] for v, o in zip(vals, ok) if o]
            if not es:
                return None, [0] * len(vals)
            E = max(es)
            res = []
            for v, o in zip(vals, ok):
                if not o:
                    res.append(0)
                    continue
                a = math.floor(Fraction(abs(v)) * Fraction(2) ** (mb - trim - E))
                res.append(-a if v < 0 else a)
            return E, res

        ra = [ints(A[i, k0:k0 + kk]) for i in range(m)]
        rb = [ints(B[k0:k0 + kk, j]) for j in range(n)]
        for i in range(m):
            for j in range(n):
                Ea, ia = ra[i]
                Eb, ib = rb[j]
                if Ea is None or Eb is None:
                    val = 0.0
                else:
                    x = sum(p * q for p, q in zip(ia, ib))
                    val = math.ldexp(float(x), Ea + Eb + 2 * trim - 2 * mb)
                if beta_eff != 0:
                    r = float(Fraction(alpha) * Fraction(val) + Fraction(beta_eff) * Fraction(out[i, j]))
                    if beta_eff * out[i, j] != float(Fraction(beta_eff) * Fraction(out[i, j])):
                        raise AssertionError("beta C is not exact in this test")
                else:
                    r = alpha * val
                out[i, j] = float(to_dtype(r))
    return out


@pytest.mark.parametrize("dtype,trim", [(np.float64, 0), (np.float64, 6), (np.float32, 0), (np.float32, 4)])
def test_matches_exact_rationals(dtype, trim):
    rng = np.random.default_rng(13 + trim)
    m, n, k = 4, 3, 40
    A = _rowcol(rng, m, k, 12, True, dtype)
    B = _rowcol(rng, k, n, 12, False, dtype)
    A[1, 3] = 0.0
    B[5, 2] = -0.0
    mb = 53 if dtype == np.float64 else 24
    for alpha, beta, kpass in ((1.0, 0.0, 0), (0.75, 0.0, 0), (1.0, 2.0, 0), (-1.5, 0.25, 16)):
        C = rng.integers(-8, 8, size=(m, n)).astype(dtype)
        out = oracle.ozaki2_gemm("N", "N", m, n, k, alpha, _fortran(A), m, _fortran(B), k, beta, _fortran(C), m,
                                 trim=trim, kpass=kpass)
        ref = _restated(A, B, alpha, beta, C, trim, mb, kpass, dtype)
        assert np.array_equal(out.astype(np.float64), _fortran(ref)), (alpha, beta, kpass)


def test_special_values_are_zero():
    """zero, subnormal, inf and nan inputs count as zero (ozaki_common.cl:789-805)"""
    A = np.array([[1.5, 5e-324, np.inf, -2.0], [np.nan, 3.0, 0.0, 1e-310]])
    B = np.array([[2.0], [7.0], [1.0], [0.5]])
    out = oracle.ozaki2_gemm("N", "N", 2, 1, 4, 1.0, _fortran(A), 2, _fortran(B), 4, 0.0, np.zeros(2), 2)
    assert np.array_equal(out, [1.5 * 2.0 - 2.0 * 0.5, 3.0 * 7.0])


# ---- truncation bound ----------------------------------------------------------------------------

@pytest.mark.parametrize("trim", [0, 10])
def test_truncation_bound_against_double_double(trim):
    """|C - AB| <= 2^(Ea+t-53) sum|b| + 2^(Eb+t-53) sum|a| + k 2^(Ea+Eb+2t-106) + u |C| per element"""
    rng = np.random.default_rng(14)
    m = n = k = 64
    A = _rowcol(rng, m, k, 15, True)
    B = _rowcol(rng, k, n, 15, False)
    a, b = _fortran(A), _fortran(B)
    out = oracle.ozaki2_gemm("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, np.zeros(m * n), m, trim=trim).reshape(n, m).T
    hi, lo = oracle.gemm_dd("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, None, m)
    ref = (hi + lo).reshape(n, m).T
    Ea = np.frexp(np.abs(A).max(axis=1))[1].astype(np.float64)[:, None]
    Eb = np.frexp(np.abs(B).max(axis=0))[1].astype(np.float64)[None, :]
    bound = (np.exp2(Ea + trim - 53) * np.abs(B).sum(axis=0)[None, :] + np.exp2(Eb + trim - 53) * np.abs(A).sum(axis=1)[:, None]
             + k * np.exp2(Ea + Eb + 2 * trim - 106)) + U * np.abs(ref)
    err = np.abs(out - ref)
    assert np.all(err <= bound * (1 + 1e-9))
    if trim == 0:  # exponents spread over 2^+-30 within rows: truncation keeps the normwise error near 1e-15
        assert np.linalg.norm(out - ref) / np.linalg.norm(ref) <= 1e-14


# ---- planted mistakes --------------------------------------------------------------------------

@pytest.mark.parametrize("plant", [oracle.OZ2_PLANT_ROUND, oracle.OZ2_PLANT_SIGNB, oracle.OZ2_PLANT_SCALE,
                                   oracle.OZ2_PLANT_FPSUM])
def test_planted_mistakes_rejected(plant):
    rng = np.random.default_rng(15)
    m, n, k = 4, 3, 40
    A = _rowcol(rng, m, k, 12, True)
    B = _rowcol(rng, k, n, 12, False)
    C = np.zeros((m, n))
    good = _fortran(_restated(A, B, 1.0, 0.0, C, 0, 53, 0, np.float64))
    bad = oracle.ozaki2_gemm("N", "N", m, n, k, 1.0, _fortran(A), m, _fortran(B), k, 0.0, _fortran(C), m, plant=plant)
    assert not np.array_equal(bad, good)
