"""Pins for the oracle (CPU only).

The reference ships no golden vectors or known-answer tests for the stack multiply (its tests/
cover LIBXS predict files and PTX retargeting), and its host path needs LIBXS, which is absent.
The oracle is therefore pinned by:
  - exact rational results (tests/golden/stack_exact.json, written by make_golden.py with
    fractions.Fraction straight from the definition) on asymmetric shapes where a transposed
    operand, a wrong index base or a dropped term changes the answer;
  - closed forms: identity A gives C + sum of B blocks, zero B leaves C unchanged;
  - planted mistakes: each `plant` mode of oracle_stack_dd must be rejected by the pins above;
  - an independent pure-Python restatement of glibc's rand() (the published additive feedback
    generator) and of INIT_STACK (samples/smm/acc_bench.h:48-79) against the C restatements;
  - the timed CPU baseline against the double-double oracle.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "stack_exact.json")


def load_cases():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def _arr(hexes, dtype):
    return np.array([float.fromhex(x) for x in hexes], dtype=dtype)


def _frac(s):
    p, q = s.split("/")
    return Fraction(int(p), int(q))


def run_case(case, plant=oracle.PLANT_NONE, b_rowmajor=False, pad=0):
    dt = np.float64 if case["dtype"] == "f64" else np.float32
    m, n, k = case["m"], case["n"], case["k"]
    a = _arr(case["a"], dt)
    b = _arr(case["b_colmajor"], dt)
    if b_rowmajor:  # present B in the kernel's layout (row-major k-by-n blocks)
        nb = b.size // (k * n)
        b = np.ascontiguousarray(b.reshape(nb, n, k).transpose(0, 2, 1).reshape(-1))
    c_in = _arr(case["c_in"], dt)
    stack = np.array(case["stack"], dtype=np.int32)
    if pad:  # room for reads shifted past the last block (planted index-base mistake)
        a = np.concatenate([a, np.zeros(pad, dtype=dt)])
        b = np.concatenate([b, np.zeros(pad, dtype=dt)])
    return oracle.stack_dd(m, n, k, stack, a, b, c_in=c_in, b_rowmajor=b_rowmajor, plant=plant)


def max_rel_dev(case, hi, lo):
    exact = [_frac(s) for s in case["c_exact"]]
    scale = max(abs(x) for x in exact) or Fraction(1)
    dev = max(abs(Fraction(float(h)) + Fraction(float(l)) - x) for h, l, x in zip(hi, lo, exact))
    return float(dev / scale)


@pytest.mark.parametrize("case", load_cases(), ids=lambda c: c["name"])
@pytest.mark.parametrize("b_rowmajor", [False, True], ids=["b_colmajor", "b_rowmajor"])
def test_exact_golden(case, b_rowmajor):
    hi, lo = run_case(case, b_rowmajor=b_rowmajor)
    # double-double carries ~106 bits: the residual against the exact value is far below 2^-90
    assert max_rel_dev(case, hi, lo) < 2.0**-90
    # and hi is the exact value rounded to double (within half an ulp)
    exact = [_frac(s) for s in case["c_exact"]]
    for h, x in zip(hi, exact):
        assert abs(Fraction(float(h)) - x) <= Fraction(np.spacing(abs(float(h))) if h != 0 else 0) / 2 + Fraction(1, 2**1074)


@pytest.mark.parametrize("plant", [oracle.PLANT_BLAYOUT, oracle.PLANT_DROPK, oracle.PLANT_BASE0, oracle.PLANT_NOACC])
def test_planted_mistakes_are_caught(plant):
    """Each plausible mistake, planted in the oracle, fails the exact fixtures."""
    caught = 0
    for case in load_cases():
        if case["name"].startswith("zero_b"):
            continue
        hi, lo = run_case(case, plant=plant, pad=1)
        if max_rel_dev(case, hi, lo) > 1e-10:
            caught += 1
    assert caught >= 3


def test_identity_closed_form():
    rng = np.random.default_rng(1)
    m = k = 6
    n = 5
    a = np.eye(m).reshape(-1, order="F").copy()  # one identity A block
    b = rng.uniform(-1, 1, size=4 * k * n)
    c0 = rng.uniform(-1, 1, size=m * n)
    stack = np.array([[1, i * k * n + 1, 1] for i in range(4)], dtype=np.int32)
    hi, lo = oracle.stack_dd(m, n, k, stack, a, b, c_in=c0)
    expect = c0 + b.reshape(4, n, k).transpose(0, 2, 1).sum(axis=0).reshape(-1, order="F")
    assert np.allclose(hi + lo, expect, rtol=0, atol=1e-15)


def test_zero_b_closed_form():
    rng = np.random.default_rng(2)
    m, n, k = 7, 3, 5
    a = rng.uniform(-1, 1, size=2 * m * k)
    b = np.zeros(k * n)
    c0 = rng.uniform(-1, 1, size=2 * m * n)
    stack = np.array([[1, 1, 1], [m * k + 1, 1, m * n + 1]], dtype=np.int32)
    hi, lo = oracle.stack_dd(m, n, k, stack, a, b, c_in=c0)
    assert np.array_equal(hi, c0) and not lo.any()


def test_numpy_cross_check():
    """Larger random stack: the oracle equals a float64 numpy evaluation within rounding."""
    rng = np.random.default_rng(3)
    m, n, k, na, nb, nc, S = 13, 7, 9, 20, 20, 10, 300
    a = rng.uniform(-1, 1, size=(na, m * k))
    b = rng.uniform(-1, 1, size=(nb, k * n))
    c0 = rng.uniform(-1, 1, size=(nc, m * n))
    ia, ib, ic = rng.integers(0, na, S), rng.integers(0, nb, S), np.sort(rng.integers(0, nc, S))
    stack = np.stack([ia * m * k + 1, ib * k * n + 1, ic * m * n + 1], axis=1).astype(np.int32)
    hi, lo = oracle.stack_dd(m, n, k, stack, a, b, c_in=c0.ravel())
    A = a.reshape(na, k, m).transpose(0, 2, 1)  # column-major blocks -> (m, k)
    B = b.reshape(nb, n, k).transpose(0, 2, 1)  # column-major blocks -> (k, n)
    C = c0.reshape(nc, n, m).transpose(0, 2, 1).copy()
    for x, y, z in zip(ia, ib, ic):
        C[z] += A[x] @ B[y]
    ref = C.transpose(0, 2, 1).reshape(-1)
    assert np.max(np.abs(ref - (hi + lo))) < 1e-13


def test_error_metric_known_values():
    ref_hi = np.array([1.0, -2.0, 0.5])
    ref_lo = np.zeros(3)
    d = 2.0**-30  # exactly representable perturbation
    test = np.array([1.0, -2.0 + d, 0.5])
    maxnorm, normwise, maxref, maxerr = oracle.error(test, ref_hi, ref_lo)
    assert maxref == 2.0 and maxerr == d
    assert maxnorm == d / 2
    assert abs(normwise - d / np.sqrt(1 + 4 + 0.25)) < 1e-25
    nan = oracle.error(np.array([np.nan, 1.0]), np.array([1.0, 1.0]), np.zeros(2))
    assert not nan[0] <= 1.0


def test_transpose_closed_form():
    rng = np.random.default_rng(4)
    m, n = 4, 9
    data = rng.uniform(-1, 1, size=5 * m * n)
    trs = np.array([0, 2 * m * n, 4 * m * n], dtype=np.int32)
    out = oracle.transpose(m, n, trs, 1, data.copy())  # offset 1: blocks 2 and 4 only
    blocks = data.reshape(5, n, m)  # column-major m-by-n == (n, m) row-major
    expect = data.reshape(5, m * n).copy()
    for i in (2, 4):
        expect[i] = blocks[i].T.reshape(-1)  # row-major m-by-n
    assert np.array_equal(out.reshape(5, m * n), expect)


def glibc_rand_py(seed, count):
    """glibc random_r TYPE_3 (the published additive feedback generator behind rand())."""
    r = [0] * (344 + count)
    r[0] = seed & 0xFFFFFFFF
    for i in range(1, 31):
        hi, lo = divmod(r[i - 1] if r[i - 1] < 2**31 else r[i - 1] - 2**32, 127773)
        word = 16807 * lo - 2836 * hi
        if word < 0:
            word += 2147483647
        r[i] = word
    for i in range(31, 34):
        r[i] = r[i - 31]
    for i in range(34, 344 + count):
        r[i] = (r[i - 31] + r[i - 3]) & 0xFFFFFFFF
    return [x >> 1 for x in r[344:]]


def test_glibc_rand():
    # srand(1): the well-known first outputs of glibc's rand()
    assert list(oracle.glibc_rand(1, 5)) == [1804289383, 846930886, 1681692777, 1714636915, 1957747793]
    assert list(oracle.glibc_rand(25071975, 2000)) == glibc_rand_py(25071975, 2000)


def init_stack_py(stack_size, rnd, mn, mk, kn, nc, na, nb):
    """INIT_STACK (samples/smm/acc_bench.h:48-79), restated independently of oracle_init_stack."""
    navg = stack_size // nc
    nimb = max(1, navg - 4)
    out, i, c, ntop = [], 0, 0, 0
    while i < stack_size:
        r = rnd[i % len(rnd)]
        nxt = c + 1
        ntop += max(1, navg + (r % (2 * nimb) - nimb))  # at least one entry (the macro loops for navg = 1)
        ntop = min(ntop, stack_size)
        while i < ntop:
            out.append((rnd[(2 * i) % len(rnd)] % na * mk + 1, rnd[(2 * i + 1) % len(rnd)] % nb * kn + 1, c * mn + 1))
            i += 1
        if nxt < nc:
            c = nxt
    return np.array(out, dtype=np.int32)


def test_init_stack_and_mat():
    """The reference benchmark's own workload: 23^3, S=30000, nc=S/16, na=nb=10*nc (acc_bench.c:239-241)."""
    S, m = 30000, 23
    rnd = oracle.glibc_rand(25071975, 30000)  # NRAND = BATCHSIZE (acc_bench.c:40-48,144-147)
    nc = S // 16
    stack = oracle.init_stack(S, rnd, m * m, m * m, m * m, nc, 10 * nc, 10 * nc)
    assert np.array_equal(stack, init_stack_py(S, list(rnd), m * m, m * m, m * m, nc, 10 * nc, 10 * nc))
    c = stack[:, 2]
    assert np.all(np.diff(c) >= 0)  # sorted by C
    assert stack[:, 0].max() <= 10 * nc * m * m and stack.min() >= 1
    fixture = [[4826068, 8417449, 1], [4426144, 541168, 1], [1884299, 5151932, 1]]  # regression fixture
    assert stack[:3].tolist() == fixture
    # INIT_MAT closed form: value (scale*seed + scale) * (index + 1)
    mat = oracle.init_mat(np.float64, 42, m, m, 0.5)
    assert np.allclose(mat, (0.5 * 42 + 0.5) * np.arange(1, m * m + 1))


def test_cpu_baseline_matches_oracle():
    from libxstream_amd.workload import make_stack_workload

    w = make_stack_workload(23, 23, 23, 600, dtype=np.float64, seed=5, c_init="uniform")
    c = w.c.copy()
    oracle.cpu_stack(23, 23, 23, w.stack, w.a, w.b, c, nrepeat=1, warmup=0, nthreads=2)
    hi, lo = oracle.stack_dd(23, 23, 23, w.stack, w.a, w.b, c_in=w.c.ravel())
    assert oracle.error(c.ravel(), hi, lo)[0] < 1e-13


def test_init_stack_invariants():
    """INIT_STACK's structure read back from the stack it made (samples/smm/acc_bench.h:48-79): every entry's
    A and B block are rnd[2i] mod na and rnd[2i+1] mod nb of the supplied table, C blocks are 0, 1, ... in
    order, each run but the last is navg + (rnd[start] mod 2 nimb) - nimb entries long (start: its first
    entry), and the last C block takes the rest. A transposed index, a wrong base or a dropped term in the
    restatement breaks one of these."""
    for S, mk, kn, mn, nc, na, nb, R in ((30000, 529, 529, 529, 1875, 18750, 18750, 30000),
                                         (5000, 88, 198, 36, 250, 777, 555, 997), (64, 4, 6, 6, 32, 3, 2, 7)):
        rnd = oracle.glibc_rand(25071975 + S, R)
        st = oracle.init_stack(S, rnd, mn, mk, kn, nc, na, nb).astype(np.int64)
        i = np.arange(S)
        assert np.array_equal((st[:, 0] - 1) // mk, rnd[(2 * i) % R] % na)
        assert np.array_equal((st[:, 1] - 1) // kn, rnd[(2 * i + 1) % R] % nb)
        assert np.all((st[:, 0] - 1) % mk == 0) and np.all((st[:, 1] - 1) % kn == 0) and np.all((st[:, 2] - 1) % mn == 0)
        cblk = (st[:, 2] - 1) // mn
        assert cblk[0] == 0 and np.all(np.diff(cblk) >= 0) and np.all(np.diff(cblk) <= 1)
        starts = np.flatnonzero(np.diff(np.concatenate([[-1], cblk])))
        lengths = np.diff(np.concatenate([starts, [S]]))
        navg = S // nc
        nimb = max(1, navg - 4)
        expect = navg + rnd[starts[:-1] % R] % (2 * nimb) - nimb
        assert np.array_equal(lengths[:-1], expect)
        assert cblk[-1] == min(nc - 1, len(starts) - 1)
    # navg = 1 (nc = S): the macro's step is 0 for an even draw and it would never end; one entry per run here
    rnd = oracle.glibc_rand(7, 50)
    st = oracle.init_stack(50, rnd, 1, 1, 1, 50, 5, 5)
    assert np.array_equal(st, init_stack_py(50, list(rnd), 1, 1, 1, 50, 5, 5))
    assert np.array_equal(st[:, 2], np.arange(1, 51))
