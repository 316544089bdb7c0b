#!/usr/bin/env python3
"""Write tests/golden/stack_exact.json: small DBCSR stacks with exact expected results.

The expected C is computed in exact rational arithmetic (fractions.Fraction) straight from the
definition C[c-1 + i + j*m] += sum_l A[a-1 + i + l*m] * B[b-1 + l + j*k] (A m-by-k and C m-by-n
column-major, B k-by-n column-major, 1-based element offsets; samples/smm/acc_bench.c:363-374), so
it pins the oracle independently of any floating-point summation order. Values are stored as
float.hex strings (inputs) and exact fractions "p/q" (outputs). Deterministic: rerunning rewrites
the same file.
"""
import json
import os
import random
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))


def exact_apply(m, n, k, stack, a, b, c_in):
    c = [Fraction(x) for x in c_in]
    for (ia, ib, ic) in stack:
        for j in range(n):
            for i in range(m):
                s = Fraction(0)
                for l in range(k):
                    s += Fraction(a[ia - 1 + i + l * m]) * Fraction(b[ib - 1 + l + j * k])
                c[ic - 1 + i + j * m] += s
    return c


def f32(x):
    import struct

    return struct.unpack("f", struct.pack("f", x))[0]


def make_case(name, rng, m, n, k, na, nb, nc, entries, dtype="f64", kind="random"):
    if kind == "identity":  # A blocks are identities (m == k): C = C + sum of the B blocks
        a = []
        for _ in range(na):
            a += [1.0 if i == l else 0.0 for l in range(k) for i in range(m)]
    else:
        a = [rng.uniform(-1, 1) for _ in range(na * m * k)]
    if kind == "zero_b":
        b = [0.0] * (nb * k * n)
    else:
        b = [rng.uniform(-1, 1) for _ in range(nb * k * n)]
    c_in = [rng.uniform(-1, 1) for _ in range(nc * m * n)]
    if dtype == "f32":
        a, b, c_in = [f32(x) for x in a], [f32(x) for x in b], [f32(x) for x in c_in]
    stack = [[ia * m * k + 1, ib * k * n + 1, ic * m * n + 1] for (ia, ib, ic) in entries]
    c = exact_apply(m, n, k, stack, a, b, c_in)
    return {
        "name": name, "dtype": dtype, "m": m, "n": n, "k": k,
        "stack": stack,
        "a": [float(x).hex() for x in a],
        "b_colmajor": [float(x).hex() for x in b],
        "c_in": [float(x).hex() for x in c_in],
        "c_exact": ["%d/%d" % (x.numerator, x.denominator) for x in c],
    }


def main():
    rng = random.Random(25071975)
    cases = []
    cases.append(make_case("tiny_2x2x2", rng, 2, 2, 2, 2, 2, 2, [(0, 1, 0), (1, 0, 0), (1, 1, 1)]))
    cases.append(make_case("asym_4x9x22", rng, 4, 9, 22, 3, 3, 2, [(0, 2, 0), (2, 1, 0), (1, 0, 1), (2, 2, 1)]))
    cases.append(make_case("asym_9x4x3", rng, 9, 4, 3, 2, 2, 2, [(1, 0, 0), (0, 1, 1), (1, 1, 1)]))
    cases.append(make_case("vec_3x1x5", rng, 3, 1, 5, 2, 2, 1, [(0, 0, 0), (1, 1, 0)]))
    cases.append(make_case("scalar_1x1x1", rng, 1, 1, 1, 2, 2, 2, [(0, 1, 1), (1, 0, 1), (1, 1, 0)]))
    cases.append(make_case("unsorted_3x3x3", rng, 3, 3, 3, 3, 3, 3,
                           [(0, 0, 2), (1, 1, 0), (2, 2, 2), (0, 1, 1), (2, 0, 0), (1, 2, 2)]))
    cases.append(make_case("identity_4x3x4", rng, 4, 3, 4, 1, 3, 1, [(0, 0, 0), (0, 1, 0), (0, 2, 0)], kind="identity"))
    cases.append(make_case("zero_b_5x5x5", rng, 5, 5, 5, 2, 1, 2, [(0, 0, 0), (1, 0, 1)], kind="zero_b"))
    cases.append(make_case("f32_4x9x22", rng, 4, 9, 22, 2, 2, 2, [(0, 1, 0), (1, 0, 1), (1, 1, 1)], dtype="f32"))
    cases.append(make_case("long_run_2x3x4", rng, 2, 3, 4, 2, 2, 1, [(i % 2, (i // 2) % 2, 0) for i in range(40)]))
    with open(os.path.join(HERE, "stack_exact.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py", "cases": cases}, f, indent=0)


if __name__ == "__main__":
    main()
