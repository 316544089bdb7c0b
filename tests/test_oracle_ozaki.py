"""Pins of the oracle's Ozaki scheme I restatement (oracle/ozaki_oracle.c), CPU only.

What fixes the restated algorithm:
  - the reference's own digit formulation (aligned mantissa, 7-bit fields; ozaki_common.cl:789-825)
    against the device formulation, digit by digit;
  - exact products: small-integer matrices are represented exactly by few slices, so the emulated
    GEMM must equal the int64 product bit for bit (every trans combination, alpha/beta, fp32);
  - a closed form that separates truncation from rounding of the digits;
  - the rigorous truncation bound against the double-double GEMM (oracle_gemm_dd), and the bound of
    the levels a trim drops;
  - planted mistakes (dropped level, wrong level weight, rounded digits, ignored transb), each of
    which one of the pins above must reject.
"""
import numpy as np
import pytest

import oracle

U = 2.0 ** -53


def _fortran(x):
    return np.asfortranarray(x).ravel(order="F")


def _op_store(X, trans):
    """storage (column-major flat) and leading dimension of X such that op(storage) = X"""
    S = X.T if trans == "T" else X
    return _fortran(S), S.shape[0]


def _int_case(rng, m, n, k, lim=1000):
    A = rng.integers(-lim, lim + 1, size=(m, k)).astype(np.float64)
    B = rng.integers(-lim, lim + 1, size=(k, n)).astype(np.float64)
    return A, B


def _rowcol(rng, m, k, e, rows):
    mant = rng.uniform(1, 2, size=(m, k)) * rng.choice([-1.0, 1.0], size=(m, k))
    ex = rng.integers(-e, e + 1, size=(m, 1) if rows else (1, k))
    return mant * np.exp2(ex)


# ---- digits ------------------------------------------------------------------------------------

def test_digits_match_reference_formulation():
    rng = np.random.default_rng(3)
    rows = [
        rng.uniform(-1, 1, 200),
        rng.uniform(1, 2, 200) * np.exp2(rng.integers(-40, 41, 200)) * rng.choice([-1, 1], 200),
        np.concatenate([[0.0, -0.0, 5e-324, 2.2e-308, np.inf, -np.inf, np.nan], rng.uniform(-3, 3, 50)]),
        np.array([1.0, 0.5, 0.25, 2.0 ** -52, 1 - 2.0 ** -53, -(1 - 2.0 ** -53)]),
    ]
    for v in rows:
        ex, d = oracle.oz_digits(v, 8)
        be, dr = oracle.oz_digits_ref(v, 8)
        finite = np.isfinite(v) & (np.abs(v) >= 2.2250738585072014e-308)
        if finite.any():
            assert ex == be - 1022  # 2^(ex-1) <= row max < 2^ex
        assert np.array_equal(d[:7], dr[:7])
        # the 8th field: the reference keeps the 4 remaining mantissa bits, the device 7 bits
        assert np.array_equal(np.sign(d[7]) * (np.abs(d[7].astype(int)) >> 3), dr[7].astype(int))
        # zero, subnormal, Inf and NaN have no digits
        assert not d[:, ~finite].any()


def test_digits_reconstruct_truncated_value():
    rng = np.random.default_rng(4)
    v = rng.uniform(-1, 1, 500) * np.exp2(rng.integers(-5, 6, 500))
    for s in range(1, 9):
        ex, d = oracle.oz_digits(v, s)
        w = np.exp2(-7.0 * np.arange(1, s + 1))
        rec = np.ldexp((d.astype(np.float64) * w[:, None]).sum(axis=0), ex)
        tail = v - rec
        # truncation toward zero: same sign (or zero) and below one unit of the last digit
        assert np.all(np.abs(tail) < np.ldexp(2.0 ** (-7 * s), ex))
        assert np.all(tail * v >= 0)
        assert np.all(np.abs(d) <= 127)


# ---- exact products ----------------------------------------------------------------------------

@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("s", [2, 3, 8])
def test_integer_products_exact(ta, tb, s):
    rng = np.random.default_rng(11 + s)
    m, n, k = 37, 29, 70
    A, B = _int_case(rng, m, n, k)
    a, lda = _op_store(A, ta)
    b, ldb = _op_store(B, tb)
    C0 = rng.integers(-50, 51, size=(m, n)).astype(np.float64)
    out = oracle.ozaki_gemm(ta, tb, m, n, k, 2.0, a, lda, b, ldb, -1.0, _fortran(C0), m, s)
    exact = 2 * (A.astype(np.int64) @ B.astype(np.int64)) - C0.astype(np.int64)
    assert np.array_equal(out, _fortran(exact).astype(np.float64))


def test_integer_products_exact_fp32_and_passes():
    rng = np.random.default_rng(5)
    m, n, k = 20, 33, 150
    A, B = _int_case(rng, m, n, k, lim=100)
    exact = _fortran(A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    out = oracle.ozaki_gemm("N", "N", m, n, k, 1.0, _fortran(A).astype(np.float32), m, _fortran(B).astype(np.float32),
                            k, 0.0, np.zeros(m * n, np.float32), m, 4)
    assert out.dtype == np.float32 and np.array_equal(out.astype(np.float64), exact)
    for maxk in (64, 100, 0):
        out = oracle.ozaki_gemm("N", "N", m, n, k, 1.0, _fortran(A), m, _fortran(B), k, 0.0, np.zeros(m * n), m, 3,
                                maxk=maxk)
        assert np.array_equal(out, exact)


def test_alpha_zero_and_k_zero():
    rng = np.random.default_rng(6)
    C0 = rng.uniform(-1, 1, 12)
    a = np.full(12, np.nan)
    assert np.array_equal(oracle.ozaki_gemm("N", "N", 3, 4, 3, 0.0, a, 3, a, 3, 2.0, C0, 3, 8), 2 * C0)
    assert np.array_equal(oracle.ozaki_gemm("N", "N", 3, 4, 0, 1.0, a, 3, a, 1, 0.0, C0, 3, 8), np.zeros(12))


def test_truncation_closed_form():
    """A = [0.99], B = [1]: one slice keeps trunc(0.99 * 128) / 128 = 126/128 (rounding would keep 127/128)."""
    out = oracle.ozaki_gemm("N", "N", 1, 1, 1, 1.0, np.array([0.99]), 1, np.array([1.0]), 1, 0.0, np.zeros(1), 1, 1)
    assert out[0] == 126.0 / 128.0
    # negative values truncate toward zero as well
    out = oracle.ozaki_gemm("N", "N", 1, 1, 1, 1.0, np.array([-0.99]), 1, np.array([1.0]), 1, 0.0, np.zeros(1), 1, 1)
    assert out[0] == -126.0 / 128.0


# ---- bounds against the double-double GEMM ----------------------------------------------------

def _bound(A, B, s, trim):
    """rigorous |C_oz - AB| bound per entry: truncation tails, dropped levels, FP64 folds"""
    k = A.shape[1]
    ea = np.array([np.frexp(np.abs(r).max())[1] for r in A])
    eb = np.array([np.frexp(np.abs(c).max())[1] for c in B.T])
    absprod = np.abs(A) @ np.abs(B)
    tail = np.exp2(ea - 7.0 * s)[:, None] * np.abs(B).sum(axis=0)[None, :] + \
        np.abs(A).sum(axis=1)[:, None] * np.exp2(eb - 7.0 * s)[None, :]
    cutoff = 2 * (s - 1) - trim
    # level L holds min(L, 2(s-1) - L) + 1 pairs, each below 127^2 k 2^(-7(L+2)) in units of 2^(ea+eb)
    dropped = sum((min(L, 2 * (s - 1) - L) + 1) * 127.0 ** 2 * k * 2.0 ** (-7 * (L + 2))
                  for L in range(cutoff + 1, 2 * s - 1))
    fold = (2 * s + 2) * U * (absprod + tail)
    return tail + fold + dropped * np.exp2(ea[:, None] + eb[None, :])


@pytest.mark.parametrize("s,trim", [(3, 0), (5, 0), (7, 0), (7, 6), (8, 0), (8, 7), (6, 3)])
def test_error_within_truncation_bound(s, trim):
    rng = np.random.default_rng(100 + s + trim)
    m, n, k = 40, 35, 130
    A = rng.uniform(-1, 1, (m, k)) * np.exp2(rng.integers(-20, 21, (m, k)))
    B = rng.uniform(-1, 1, (k, n)) * np.exp2(rng.integers(-20, 21, (k, n)))
    a, b = _fortran(A), _fortran(B)
    out = oracle.ozaki_gemm("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, np.zeros(m * n), m, s, trim=trim)
    hi, lo = oracle.gemm_dd("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, None, m)
    err = np.abs(out - hi - lo).reshape(n, m).T
    assert np.all(err <= _bound(A, B, s, trim))


def test_error_decreases_with_slices():
    rng = np.random.default_rng(8)
    m = n = k = 96
    A, B = _rowcol(rng, m, k, 10, True), _rowcol(rng, k, n, 10, False)
    a, b = _fortran(A), _fortran(B)
    hi, lo = oracle.gemm_dd("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, None, m)
    ref = hi + lo
    errs = [np.linalg.norm(oracle.ozaki_gemm("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, np.zeros(m * n), m, s) - ref)
            / np.linalg.norm(ref) for s in range(2, 9)]
    assert all(e1 < e0 / 50 for e0, e1 in zip(errs, errs[1:-1]))
    # per-row/column exponents: 7 slices over the triangle reach 1e-14, 8 slices reach FP64
    e7t = np.linalg.norm(oracle.ozaki_gemm("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, np.zeros(m * n), m, 7, trim=6)
                         - ref) / np.linalg.norm(ref)
    assert e7t <= 1e-14 and errs[-1] <= 1e-15


# ---- planted mistakes ---------------------------------------------------------------------------

@pytest.mark.parametrize("plant", [oracle.OZ_PLANT_DROPLEVEL, oracle.OZ_PLANT_WEIGHT, oracle.OZ_PLANT_TRANSB,
                                   oracle.OZ_PLANT_ROUND])
def test_planted_mistakes_are_caught(plant):
    rng = np.random.default_rng(9)
    m, n, k = 24, 20, 40
    A, B = _int_case(rng, m, n, k)
    a, lda = _op_store(A, "N")
    b, ldb = _op_store(B, "T")
    exact = _fortran(A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    out = oracle.ozaki_gemm("N", "T", m, n, k, 1.0, a, lda, b, ldb, 0.0, np.zeros(m * n), m, 3, plant=plant)
    one = oracle.ozaki_gemm("N", "N", 1, 1, 1, 1.0, np.array([0.99]), 1, np.array([1.0]), 1, 0.0, np.zeros(1), 1, 1,
                            plant=plant)
    caught_exact = not np.array_equal(out, exact)
    caught_closed = one[0] != 126.0 / 128.0
    assert caught_exact or caught_closed
