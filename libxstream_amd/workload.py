"""Seeded synthetic DBCSR stack workloads (pools of A/B/C blocks and a sorted parameter stack).

The construction follows the benchmark driver of the reference (samples/smm/acc_bench.c:239-278,
INIT_STACK in samples/smm/acc_bench.h:48-79): C blocks are visited in increasing order, each C-run
has a pseudo-random length around an average, and a/b block indices are uniform. Values are
uniform[-1, 1] from a numpy PCG64 seed instead of INIT_MAT's arithmetic progression.

Offsets are 1-based element offsets (idx * blocksize + 1); A and C blocks are column-major;
B blocks are generated column-major k-by-n (the host view) and ``b_device_layout`` gives the
row-major k-by-n layout the stack kernel consumes (after libsmm_acc_transpose).
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class StackWorkload:
    m: int
    n: int
    k: int
    dtype: type
    stack: np.ndarray  # int32, (S, 3): 1-based element offsets (a, b, c)
    a: np.ndarray      # (na, m*k) column-major blocks
    b: np.ndarray      # (nb, k*n) column-major blocks (host view)
    c: np.ndarray      # (nc, m*n) initial C blocks
    nc: int
    na: int
    nb: int

    @property
    def stack_size(self):
        return int(self.stack.shape[0])

    def b_device_layout(self):
        """B pool with every block stored row-major k-by-n (what libsmm_acc_transpose produces)."""
        nb, k, n = self.b.shape[0], self.k, self.n
        return np.ascontiguousarray(self.b.reshape(nb, n, k).transpose(0, 2, 1).reshape(nb, k * n))

    def flops(self):
        return 2.0 * self.stack_size * self.m * self.n * self.k

    def bytes_alg(self):
        """samples/smm/smm_kernel.c:504-506: S * ts * (mk + kn + 2mn)"""
        ts = np.dtype(self.dtype).itemsize
        m, n, k = self.m, self.n, self.k
        return float(self.stack_size) * ts * (m * k + k * n + 2 * m * n)


def run_lengths(rng, stack_size, nc, spread):
    """Run lengths for C blocks 0..nc-1 around stack_size/nc (+-spread); the last run absorbs the rest."""
    avg = stack_size / nc
    lo = max(1, int(round(avg)) - spread)
    hi = max(lo, int(round(avg)) + spread)
    lengths = rng.integers(lo, hi + 1, size=nc)
    cidx = np.repeat(np.arange(nc, dtype=np.int64), lengths)
    if cidx.size >= stack_size:
        cidx = cidx[:stack_size]
    else:
        cidx = np.concatenate([cidx, np.full(stack_size - cidx.size, nc - 1, dtype=np.int64)])
    return cidx


def make_stack_workload(m, n, k, stack_size, nc=None, na=None, nb=None, dtype=np.float64, seed=25071975,
                        spread=4, c_init="zero"):
    """Sorted-by-C stack over seeded pools.

    Defaults keep the proportions of the 23^3 stack workload (2000 A, 2000 B, 1000 C blocks for
    30000 entries): nc = S/30, na = nb = 2*nc, run lengths 30 +- 4.
    """
    rng = np.random.default_rng(seed)
    nc = max(1, int(round(stack_size / 30.0))) if nc is None else int(nc)
    nc = min(nc, stack_size) if stack_size > 0 else nc
    na = 2 * nc if na is None else int(na)
    nb = 2 * nc if nb is None else int(nb)
    mk, kn, mn = m * k, k * n, m * n
    if stack_size > 0:
        cidx = run_lengths(rng, stack_size, nc, spread)
        aidx = rng.integers(0, na, size=stack_size)
        bidx = rng.integers(0, nb, size=stack_size)
        stack = np.empty((stack_size, 3), dtype=np.int64)
        stack[:, 0] = aidx * mk + 1
        stack[:, 1] = bidx * kn + 1
        stack[:, 2] = cidx * mn + 1
        if stack.max() >= 2**31:
            raise ValueError("offsets exceed int32")
        stack = stack.astype(np.int32)
    else:
        stack = np.zeros((0, 3), dtype=np.int32)
    a = rng.uniform(-1.0, 1.0, size=(na, mk)).astype(dtype)
    b = rng.uniform(-1.0, 1.0, size=(nb, kn)).astype(dtype)
    if c_init == "zero":
        c = np.zeros((nc, mn), dtype=dtype)
    else:
        c = rng.uniform(-1.0, 1.0, size=(nc, mn)).astype(dtype)
    return StackWorkload(m, n, k, dtype, stack, a, b, c, nc, na, nb)


def mixed_shapes(count, seed, dim_range=(4, 64), len_range=(1000, 200000)):
    """(m, n, k, stack_size) drawn log-uniformly: the mixed-shape batch of 24 stacks uses the defaults."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(dim_range[0]), np.log(dim_range[1])
    dims = np.rint(np.exp(rng.uniform(lo, hi, size=(count, 3)))).astype(int)
    lens = np.rint(np.exp(rng.uniform(np.log(len_range[0]), np.log(len_range[1]), size=count))).astype(int)
    return [(int(d[0]), int(d[1]), int(d[2]), int(s)) for d, s in zip(dims, lens)]


def make_mixed_batch(count=24, seed=25071975, dtype=np.float64, dim_range=(4, 64), len_range=(1000, 200000),
                     c_init="zero"):
    """Stack workloads for a mixed-shape batch: shapes and lengths from mixed_shapes(), pools per stack."""
    return [make_stack_workload(m, n, k, s, dtype=dtype, seed=seed + 7 * i, c_init=c_init)
            for i, (m, n, k, s) in enumerate(mixed_shapes(count, seed, dim_range, len_range))]
