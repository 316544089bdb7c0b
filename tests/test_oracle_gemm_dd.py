"""Pins of the double-double GEMM oracle (oracle.gemm_dd): C = alpha op(A) op(B) + beta C, column-major.

It is the reference of the Ozaki tests, so every path is pinned against exact arithmetic:
  - integer operands and dyadic alpha/beta: every product and sum is exact in double, so hi must equal
    the exact integer result and lo must be zero, for every transpose combination and alpha/beta;
  - arbitrary doubles and non-dyadic alpha on a small case: hi + lo within 2^-100 of the exact rational
    value (fractions.Fraction);
  - the sampled-entry mode returns the same (hi, lo) as the full mode at the listed entries;
  - planted mistakes (the other transpose of a non-square operand, a dropped beta term) are rejected.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle


def _store(X, trans):
    S = X.T if trans in ("T", "C") else X
    return np.asfortranarray(S).ravel(order="F"), max(1, S.shape[0])


@pytest.mark.parametrize("ta", ["N", "T", "C"])
@pytest.mark.parametrize("tb", ["N", "T", "C"])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (2.5, 1.0), (-0.75, -3.0), (0.0, 0.5)])
def test_exact_integer_products(ta, tb, alpha, beta):
    m, n, k = 7, 5, 11
    rng = np.random.default_rng(abs(hash((ta, tb, alpha, beta))) % 2**31)
    A = rng.integers(-64, 65, (m, k)).astype(np.float64)
    B = rng.integers(-64, 65, (k, n)).astype(np.float64)
    C = rng.integers(-1000, 1001, (m, n)).astype(np.float64)
    a, lda = _store(A, ta)
    b, ldb = _store(B, tb)
    c = np.asfortranarray(C).ravel(order="F")
    hi, lo = oracle.gemm_dd(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, m)
    exact = alpha * (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64) + beta * C  # exact: small dyadics
    assert np.array_equal(hi, np.asfortranarray(exact).ravel(order="F"))
    assert not lo.any()


def test_rational_accuracy_and_sampling():
    m, n, k = 3, 4, 5
    rng = np.random.default_rng(11)
    A, B, C = rng.uniform(-1, 1, (m, k)), rng.uniform(-1, 1, (k, n)), rng.uniform(-1, 1, (m, n))
    alpha, beta = 1.0 / 3.0, -0.1
    for ta, tb in (("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")):
        a, lda = _store(A, ta)
        b, ldb = _store(B, tb)
        c = np.asfortranarray(C).ravel(order="F")
        hi, lo = oracle.gemm_dd(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, m)
        for j in range(n):
            for i in range(m):
                ref = Fraction(alpha) * sum(Fraction(A[i, l]) * Fraction(B[l, j]) for l in range(k)) + \
                    Fraction(beta) * Fraction(C[i, j])
                got = Fraction(hi[i + j * m]) + Fraction(lo[i + j * m])
                assert abs(got - ref) <= abs(ref) * Fraction(1, 2**100) + Fraction(1, 2**1000), (ta, tb, i, j)
        rows = np.array([2, 0, 1, 2], dtype=np.int32)
        cols = np.array([3, 0, 2, 1], dtype=np.int32)
        shi, slo = oracle.gemm_dd(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, m, rows=rows, cols=cols)
        assert np.array_equal(shi, hi[rows + cols * m]) and np.array_equal(slo, lo[rows + cols * m])


def test_planted_mistakes_rejected():
    m, n, k = 6, 6, 6  # square, so a wrong transpose still has valid strides
    rng = np.random.default_rng(5)
    A = rng.integers(-9, 10, (m, k)).astype(np.float64)
    B = rng.integers(-9, 10, (k, n)).astype(np.float64)
    C = rng.integers(-9, 10, (m, n)).astype(np.float64)
    exact = np.asfortranarray(2.0 * (A @ B) + 3.0 * C).ravel(order="F")
    a, _ = _store(A, "T")
    b, _ = _store(B, "N")
    c = np.asfortranarray(C).ravel(order="F")
    hi, _ = oracle.gemm_dd("T", "N", m, n, k, 2.0, a, m, b, k, 3.0, c, m)
    assert np.array_equal(hi, exact)
    wrong, _ = oracle.gemm_dd("N", "N", m, n, k, 2.0, a, m, b, k, 3.0, c, m)  # transpose ignored
    assert not np.array_equal(wrong, exact)
    no_beta, _ = oracle.gemm_dd("T", "N", m, n, k, 2.0, a, m, b, k, 0.0, c, m)
    assert not np.array_equal(no_beta, exact)
