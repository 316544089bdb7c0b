"""Pins of the oracle's Ozaki stack restatement (oracle.ozaki_stack), CPU only.

The stack oracle composes the pinned scheme I GEMM oracle entry by entry in stack order. What fixes the
composition (block addressing, B's kernel layout, the order of the C updates):
  - integer blocks: every product is exact, so the result equals the integer stack product;
  - the double-double stack oracle (oracle_stack_dd): 8 slices (all pairs or the triangle) stay at FP64
    level, the 7-slice triangle about 7 bits above it;
  - order: two entries into one C block give C + v1 + v2 in stack order (checked against a manual
    fold of two single-entry stacks).
"""
import numpy as np
import pytest

import oracle
from libxstream_amd.workload import make_stack_workload


def _int_pools(rng, m, n, k, na, nb, nc, S, lim=60):
    A = rng.integers(-lim, lim + 1, (na, m * k)).astype(np.float64)
    B = rng.integers(-lim, lim + 1, (nb, k * n)).astype(np.float64)  # kernel layout: k x n row-major blocks
    C = rng.integers(-9, 10, (nc, m * n)).astype(np.float64)
    st = np.stack([rng.integers(0, na, S) * m * k + 1, rng.integers(0, nb, S) * k * n + 1,
                   np.sort(rng.integers(0, nc, S)) * m * n + 1], 1).astype(np.int32)
    return A, B, C, st


def _int_reference(m, n, k, st, A, B, C):
    ref = C.reshape(-1).astype(np.int64).copy()
    a, b = A.reshape(-1).astype(np.int64), B.reshape(-1).astype(np.int64)
    for ao, bo, co in st:
        Ab = a[ao - 1:ao - 1 + m * k].reshape(k, m).T
        Bb = b[bo - 1:bo - 1 + k * n].reshape(k, n)
        ref[co - 1:co - 1 + m * n] += (Ab @ Bb).T.reshape(-1)
    return ref.astype(np.float64)


@pytest.mark.parametrize("shape", [(4, 9, 22), (13, 13, 13), (64, 64, 64)])
def test_integer_stack_exact(shape):
    m, n, k = shape
    rng = np.random.default_rng(sum(shape))
    A, B, C, st = _int_pools(rng, m, n, k, 6, 5, 3, 12)
    out = oracle.ozaki_stack(m, n, k, st, A, B, C, 3)  # |values| < 2^6: three 7-bit digits hold them exactly
    assert np.array_equal(out, _int_reference(m, n, k, st, A, B, C))


def test_stack_against_double_double():
    m = n = k = 23
    w = make_stack_workload(m, n, k, 90, dtype=np.float64)
    bl = w.b_device_layout()
    hi, lo = oracle.stack_dd(m, n, k, w.stack, w.a, bl, w.c, b_rowmajor=True)
    ref = (hi + lo).reshape(-1)
    err = {}
    for s, trim in ((8, 0), (8, 7), (7, 6)):
        out = oracle.ozaki_stack(m, n, k, w.stack, w.a, bl, w.c, s, trim=trim)
        err[(s, trim)] = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    # 8 slices reach FP64 level over all pairs and over the triangle; 7 over the triangle drop ~7 bits
    assert err[(8, 0)] <= 1e-15 and err[(8, 7)] <= 1e-15
    assert 10 * err[(8, 7)] < err[(7, 6)] <= 3e-14


def test_updates_in_stack_order():
    m, n, k = 3, 2, 4
    rng = np.random.default_rng(5)
    A = rng.uniform(-1, 1, (2, m * k))
    B = rng.uniform(-1, 1, (2, k * n))
    C = rng.uniform(-1, 1, (1, m * n))
    st = np.array([[1, 1, 1], [m * k + 1, k * n + 1, 1]], dtype=np.int32)
    both = oracle.ozaki_stack(m, n, k, st, A, B, C, 6)
    first = oracle.ozaki_stack(m, n, k, st[:1], A, B, C, 6)
    folded = oracle.ozaki_stack(m, n, k, st[1:], A, B, first, 6)
    assert np.array_equal(both, folded)
