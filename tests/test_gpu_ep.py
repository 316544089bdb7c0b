"""GPU parity of the entry-parallel matrix-core stack kernels (variant 2: several wavefronts share each
C-run, their partial blocks added in a fixed order through LDS; variant 3: the same on 4x4 blocks, FP64) through
the C ABI, against the oracle.

Covered: every compiled variant-2 and variant-3 kernel on its shape, FP64 and FP32, sorted stacks with ragged runs,
long runs cut into pieces (atomic flush of the pieces), stacks whose runs are interleaved (atomic flush),
the read-modify-write flush on sorted stacks, and bit-reproducibility: runs that are not cut give the same
bits on every call, for every bs, and under both flush modes.
"""
import numpy as np
import pytest

import libxstream_amd as lx
import oracle
from libxstream_amd.workload import make_stack_workload
from tests.helpers import oracle_error, run_stack, tol

pytestmark = pytest.mark.gpu

SHAPES = [(5, 5, 5), (13, 13, 13), (23, 23, 23), (32, 32, 32), (64, 64, 64), (4, 9, 22), (45, 45, 13)]
DTYPES = [np.float64, np.float32]


@pytest.fixture(scope="module")
def stream():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    s = lx.Stream("ep")
    yield s
    s.close()
    lx.finalize()


def _ep_configs(tid, m, n, k):
    """every entry-parallel kernel for the shape: 16x16 tiles (variant 2) and 4x4 blocks (variant 3, FP64, and
    variant 4, whose instructions split 16 of k over their blocks)"""
    out = []
    for variant, tm, tn, nt, kt in lx.kernel_candidates(tid, m, n, k):
        if variant in (2, 3, 4):
            out.append(lx.SmmConfig(variant=variant, tm=tm, tn=tn, nt=nt, kt=kt, bs=32, flush=2))
    return out


def _with_runs(w, lengths):
    stack = w.stack.copy()
    stack[:, 2] = np.repeat(np.arange(len(lengths), dtype=np.int32), lengths) * (w.m * w.n) + 1
    return stack


def test_every_workload_shape_has_ep_kernels():
    lx.init()
    try:
        for tid in (lx.REAL_8, lx.REAL_4):
            for shape in SHAPES:
                cfgs = _ep_configs(tid, *shape)
                assert cfgs, shape
                assert all(lx.config_valid(tid, *shape, c) for c in cfgs)
    finally:
        lx.finalize()


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
@pytest.mark.parametrize("shape", SHAPES, ids=["%dx%dx%d" % s for s in SHAPES])
def test_ep_parity(stream, shape, dtype):
    m, n, k = shape
    tid = lx.typeid(dtype)
    w = make_stack_workload(m, n, k, 3000, dtype=dtype, seed=sum(shape) * 7 + 1, c_init="uniform", spread=12)
    long_runs = [700, 3, 7, 40, 1, 64, 65, 33, 300] + [17] * 40 + [1107]
    wl = make_stack_workload(m, n, k, sum(long_runs), nc=len(long_runs), dtype=dtype, seed=sum(shape) * 7 + 2,
                             c_init="uniform")
    stack_long = _with_runs(wl, long_runs)
    rng = np.random.default_rng(sum(shape))
    stack_unsorted = np.ascontiguousarray(w.stack[rng.permutation(w.stack_size)])
    ref_sorted = oracle.stack_dd(m, n, k, w.stack, w.a, w.b, c_in=w.c.ravel())
    ref_long = oracle.stack_dd(m, n, k, stack_long, wl.a, wl.b, c_in=wl.c.ravel())
    ref_unsorted = oracle.stack_dd(m, n, k, stack_unsorted, w.a, w.b, c_in=w.c.ravel())
    for cfg in _ep_configs(tid, m, n, k):
        for bs in (1, 8, 32, 100):
            cfg.bs, cfg.s = bs, 0
            for flush in (1, 2):
                cfg.flush = flush
                out = run_stack(stream, w, config=cfg, device_transpose=False)
                assert oracle.error(out.ravel(), *ref_sorted)[0] <= tol(dtype), (cfg, "sorted")
                out = run_stack(stream, wl, config=cfg, device_transpose=False, stack=stack_long)
                assert oracle.error(out.ravel(), *ref_long)[0] <= tol(dtype), (cfg, "long runs")
            cfg.flush = 2
            out = run_stack(stream, w, config=cfg, device_transpose=False, stack=stack_unsorted)
            assert oracle.error(out.ravel(), *ref_unsorted)[0] <= tol(dtype), (cfg, "unsorted")


@pytest.mark.parametrize("shape", [(23, 23, 23), (13, 13, 13), (64, 64, 64)], ids=lambda s: "%dx%dx%d" % s)
def test_ep_bit_reproducible(stream, shape):
    """runs of at most 60 entries, never cut (split x bs >= 64): one flush per C block, the partial blocks
    summed in wavefront order, so every call, every bs >= 32 and both flush modes give the same bits"""
    m, n, k = shape
    tid = lx.REAL_8
    # explicit run lengths in 5..60 (the generator's last C block absorbs the remainder of the stack, which
    # can make that one run long enough to be cut)
    lengths = np.random.default_rng(102).integers(5, 61, size=200)
    w = make_stack_workload(m, n, k, int(lengths.sum()), nc=len(lengths), dtype=np.float64, seed=101,
                            c_init="uniform")
    stack = _with_runs(w, lengths)
    for cfg in _ep_configs(tid, m, n, k):
        cfg.s, cfg.split = 0, 2
        ref = None
        for bs, flush in ((32, 2), (32, 2), (64, 2), (32, 1), (100, 1)):
            cfg.bs, cfg.flush = bs, flush
            out = run_stack(stream, w, config=cfg, device_transpose=False, stack=stack)
            if ref is None:
                ref = out
                assert oracle_error(w, out, stack=stack)[0] <= 1e-13
            assert np.array_equal(ref.view(np.uint64), out.view(np.uint64)), (cfg, bs, flush)


# (19, 22, 6), (9, 25, 7), (31, 23, 40) and (17, 29, 11) put the last row or column of a block at the start of a
# 4x4-block kernel's element pair (the pair is shifted back by one element)
GENERIC_SHAPES = [(17, 29, 11), (31, 23, 40), (6, 21, 16), (40, 33, 9), (19, 22, 6), (9, 25, 7), (7, 10, 30), (14, 3, 9),
                  (16, 16, 5)]


def test_4x4_block_kernels_cover_their_shapes():
    """variant 3 (FP64 only) serves every shape up to 32 x 32 with exactly the blocks it needs, and no FP32 shape;
    variant 4 (FP64 and FP32) every shape up to 16 x 16"""
    lx.init()
    try:
        for m, n in ((5, 5), (23, 23), (13, 13), (4, 9), (9, 25), (32, 32), (1, 1), (2, 17)):
            v3 = [c for c in lx.kernel_candidates(lx.REAL_8, m, n, 7) if c[0] == 3]
            assert v3 and all((c[1], c[2]) == ((m + 7) // 8, (n + 7) // 8) for c in v3), (m, n, v3)
            assert not [c for c in lx.kernel_candidates(lx.REAL_4, m, n, 7) if c[0] == 3]
        assert not [c for c in lx.kernel_candidates(lx.REAL_8, 33, 8, 7) if c[0] == 3]
        # variant 4 (4x4 blocks splitting k): exactly ti = ceil(m / 4), tj = ceil(n / 4), up to 16 x 16
        for m, n in ((5, 5), (7, 10), (14, 3), (16, 16), (1, 1)):
            v4 = [c for c in lx.kernel_candidates(lx.REAL_8, m, n, 7) if c[0] == 4]
            assert v4 and all((c[1], c[2]) == ((m + 3) // 4, (n + 3) // 4) for c in v4), (m, n, v4)
            v4 = [c for c in lx.kernel_candidates(lx.REAL_4, m, n, 7) if c[0] == 4]
            assert v4 and all((c[1], c[2]) == ((m + 3) // 4, (n + 3) // 4) for c in v4), (m, n, v4)
        assert not [c for c in lx.kernel_candidates(lx.REAL_8, 17, 8, 7) if c[0] == 4]
    finally:
        lx.finalize()


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
@pytest.mark.parametrize("shape", GENERIC_SHAPES, ids=["%dx%dx%d" % s for s in GENERIC_SHAPES])
def test_ep_runtime_shape_parity(stream, shape, dtype):
    """The runtime-shape entry-parallel kernels (no kernel compiled for the shape itself): every one that fits
    the shape, on sorted stacks with ragged runs, long runs cut into pieces and unsorted stacks"""
    m, n, k = shape
    tid = lx.typeid(dtype)
    cfgs = [c for c in _ep_configs(tid, m, n, k)]
    assert cfgs, shape
    w = make_stack_workload(m, n, k, 2500, dtype=dtype, seed=sum(shape) * 11 + 1, c_init="uniform", spread=12)
    long_runs = [400, 3, 70, 1, 129, 65] + [17] * 20 + [600]
    wl = make_stack_workload(m, n, k, sum(long_runs), nc=len(long_runs), dtype=dtype, seed=sum(shape) * 11 + 2,
                             c_init="uniform")
    stack_long = _with_runs(wl, long_runs)
    stack_unsorted = np.ascontiguousarray(w.stack[np.random.default_rng(sum(shape)).permutation(w.stack_size)])
    ref_sorted = oracle.stack_dd(m, n, k, w.stack, w.a, w.b, c_in=w.c.ravel())
    ref_long = oracle.stack_dd(m, n, k, stack_long, wl.a, wl.b, c_in=wl.c.ravel())
    ref_unsorted = oracle.stack_dd(m, n, k, stack_unsorted, w.a, w.b, c_in=w.c.ravel())
    for cfg in cfgs:
        for bs in (1, 32, 100):
            cfg.bs, cfg.s = bs, 0
            for flush in (1, 2):
                cfg.flush = flush
                out = run_stack(stream, w, config=cfg, device_transpose=False)
                assert oracle.error(out.ravel(), *ref_sorted)[0] <= tol(dtype), (cfg, "sorted")
                out = run_stack(stream, wl, config=cfg, device_transpose=False, stack=stack_long)
                assert oracle.error(out.ravel(), *ref_long)[0] <= tol(dtype), (cfg, "long runs")
            cfg.flush = 2
            out = run_stack(stream, w, config=cfg, device_transpose=False, stack=stack_unsorted)
            assert oracle.error(out.ravel(), *ref_unsorted)[0] <= tol(dtype), (cfg, "unsorted")
