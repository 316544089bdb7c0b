"""GPU parity: the HIP stack kernel and transpose, through the C ABI, against the oracle.

Bar: max normalised error <= 1e-13 (FP64) / 1e-5 (FP32) against the double-double oracle
(samples/smm/acc_bench.c:357-418 host path, checked with an explicit metric in place of
libxs_matdiff_epsilon); bit-exact for the transpose (pure data movement).
"""
import zlib

import numpy as np
import pytest

import libxstream_amd as lx
import oracle
from libxstream_amd.workload import make_stack_workload
from tests.helpers import checksum, cut_limit, oracle_error, run_stack, sampled_oracle_error, tol, uncut_blocks

pytestmark = pytest.mark.gpu

SHAPES = [(5, 5, 5), (13, 13, 13), (23, 23, 23), (32, 32, 32), (64, 64, 64), (4, 9, 22), (45, 45, 13)]
DTYPES = [np.float64, np.float32]


@pytest.fixture(scope="module")
def stream():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    s = lx.Stream("test")
    yield s
    s.close()
    lx.finalize()


def _seed(*key):
    return zlib.crc32(repr(key).encode()) % (2**31)


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
@pytest.mark.parametrize("shape", SHAPES, ids=["%dx%dx%d" % s for s in SHAPES])
def test_process_parity(stream, shape, dtype):
    m, n, k = shape
    w = make_stack_workload(m, n, k, 3000, dtype=dtype, seed=_seed(shape, 1), c_init="uniform")
    out = run_stack(stream, w)
    err = oracle_error(w, out)
    assert err[0] <= tol(dtype), "max normalised error %g" % err[0]


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
def test_every_variant(stream, dtype):
    """Every compiled (tm, tn, nt) kernel on shapes that exercise padding, groups and k-slices."""
    tid = lx.typeid(dtype)
    variants = lx.kernel_variants(tid)
    assert len(variants) >= 10
    checked = 0
    for shape in [(23, 23, 23), (7, 11, 3), (16, 16, 40), (9, 4, 22)]:
        m, n, k = shape
        w = make_stack_workload(m, n, k, 700, dtype=dtype, seed=_seed(shape, 2), c_init="uniform")
        for variant, tm, tn, nt, vkt in variants:
            for kt in ((0, 8) if variant == 0 else (vkt,)):
                cfg = lx.SmmConfig(variant=variant, tm=tm, tn=tn, nt=nt, bs=16, kt=kt, flush=0)
                if not lx.config_valid(tid, m, n, k, cfg):
                    continue
                out = run_stack(stream, w, config=cfg, device_transpose=False)
                err = oracle_error(w, out)
                assert err[0] <= tol(dtype), "%s v=%d tm=%d tn=%d nt=%d kt=%d: %g" % (
                    shape, variant, tm, tn, nt, kt, err[0])
                checked += 1
    assert checked >= 40


@pytest.mark.parametrize("bs", [1, 7, 64, 1000])
def test_chunk_sizes(stream, bs):
    """Run ownership across work-group chunks: runs start and end anywhere relative to bs."""
    m = n = k = 13
    w = make_stack_workload(m, n, k, 4000, dtype=np.float64, seed=_seed("bs", bs), c_init="uniform", spread=12)
    cfg, _ = lx.config_get(lx.REAL_8, m, n, k)
    cfg.bs = bs
    cfg.s = 0  # launch with exactly this bs (no scaling for the short stack)
    out = run_stack(stream, w, config=cfg)
    assert oracle_error(w, out)[0] <= 1e-13


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
@pytest.mark.parametrize("shape", [(23, 23, 23), (5, 5, 5), (4, 9, 22), (45, 45, 13)], ids=lambda s: "%dx%dx%d" % s)
def test_short_stack_bs_scaling(stream, shape, dtype):
    """A stack shorter than the configuration's S launches with bs scaled down (smm_kernel.c:463-473):
    more work-groups, and runs longer than twice the scaled bs are cut into pieces; the same C within
    the parity bar."""
    m, n, k = shape
    w = make_stack_workload(m, n, k, 1900, dtype=dtype, seed=_seed("short", shape), c_init="uniform")
    tid = lx.typeid(dtype)
    cfg, _ = lx.config_get(tid, m, n, k)
    cfg.bs = 128
    cfg.s = 30000  # scaled: ceil(1900 * 128 / 29999) = 9 entries per work-group
    scaled = run_stack(stream, w, config=cfg, device_transpose=False)
    cfg.s = 0
    plain = run_stack(stream, w, config=cfg, device_transpose=False)
    assert oracle_error(w, scaled)[0] <= tol(dtype)
    assert oracle_error(w, plain)[0] <= tol(dtype)


@pytest.mark.parametrize("flush", [1, 2])
def test_flush_modes_bit_identical(stream, flush):
    """Sorted stack whose runs are never cut (bs 16 unscaled: runs of up to 32 entries stay whole): the
    read-modify-write and atomic flush give the same bits (one add per element)."""
    m = n = k = 23
    w = make_stack_workload(m, n, k, 6000, dtype=np.float64, seed=_seed("flush"), c_init="uniform", spread=2)
    cfg, _ = lx.config_get(lx.REAL_8, m, n, k)
    cfg.bs, cfg.s, cfg.split, cfg.bs_short = 16, 0, 2, 0
    cfg.flush = 1
    ref = run_stack(stream, w, config=cfg)
    cfg.flush = flush
    out = run_stack(stream, w, config=cfg)
    short = uncut_blocks(w, limit=32)  # the generator's tail run may be long enough to be cut
    assert np.array_equal(ref[short].view(np.uint64), out[short].view(np.uint64))
    assert oracle_error(w, out)[0] <= 1e-13


def test_deterministic(stream):
    """Runs of at most 2 bs entries are never cut, so each C block is flushed once: the same bits every time."""
    m = n = k = 32
    w = make_stack_workload(m, n, k, 5000, dtype=np.float64, seed=_seed("det"), c_init="uniform", spread=2)
    cfg, _ = lx.config_get(lx.REAL_8, m, n, k)
    cfg.bs, cfg.s, cfg.split, cfg.bs_short = 16, 0, 2, 0
    a = run_stack(stream, w, config=cfg)
    b = run_stack(stream, w, config=cfg)
    short = uncut_blocks(w, limit=32)
    assert np.array_equal(a[short].view(np.uint64), b[short].view(np.uint64))
    assert oracle_error(w, a)[0] <= 1e-13 and oracle_error(w, b)[0] <= 1e-13


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
def test_unsorted_stack(stream, dtype):
    """Interleaved C-runs (legal for the API): the default flush stays correct."""
    m, n, k = 23, 23, 23
    w = make_stack_workload(m, n, k, 3000, dtype=dtype, seed=_seed("unsorted"), c_init="uniform")
    perm = np.random.default_rng(3).permutation(w.stack_size)
    stack = np.ascontiguousarray(w.stack[perm])
    out = run_stack(stream, w, stack=stack)
    assert oracle_error(w, out, stack=stack)[0] <= tol(dtype)


UNSORTED_SHAPES = [(45, 45, 13), (23, 23, 23), (64, 64, 64), (5, 5, 5), (4, 9, 22), (13, 13, 13), (32, 32, 32)]


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
@pytest.mark.parametrize("shape", UNSORTED_SHAPES, ids=["%dx%dx%d" % s for s in UNSORTED_SHAPES])
def test_unsorted_default_path(stream, shape, dtype):
    """libsmm_acc_process with the shipped configuration of every workload shape on stacks whose C-runs
    are interleaved: several work-groups own runs of one C block, and the default flush (atomic adds,
    samples/smm/kernels/multiply.cl:657,687) keeps every update."""
    m, n, k = shape
    w = make_stack_workload(m, n, k, 3000, dtype=dtype, seed=_seed("unsorted-default", shape), c_init="uniform")
    # shuffle whole runs, then split every run in two halves placed apart: each C block has two runs
    rng = np.random.default_rng(_seed("perm", shape))
    half = [w.stack[i::2] for i in range(2)]
    stack = np.ascontiguousarray(np.concatenate([h[rng.permutation(len(h))] for h in half]))
    out = run_stack(stream, w, stack=stack)
    assert oracle_error(w, out, stack=stack)[0] <= tol(dtype)


# shapes without a tuned row whose exhaustive sweep winner was a read-modify-write (flush 1) configuration:
# the learned predictor serves them; unsorted stacks must still be correct
UNTUNED_RMW_SHAPES = [(lx.REAL_8, 30, 60, 67), (lx.REAL_8, 14, 10, 32), (lx.REAL_4, 79, 24, 36), (lx.REAL_4, 47, 57, 8)]


@pytest.mark.parametrize("key", UNTUNED_RMW_SHAPES, ids=["%d-%dx%dx%d" % s for s in UNTUNED_RMW_SHAPES])
def test_unsorted_learned_shapes(stream, key):
    tid, m, n, k = key
    cfg, src = lx.config_get(tid, m, n, k)
    assert src != 1, "tuned row for an untuned test shape"
    assert cfg.flush != 1, "a predicted configuration without atomics: %s" % cfg
    dtype = lx.nptype(tid)
    w = make_stack_workload(m, n, k, 2500, dtype=dtype, seed=_seed("unsorted-learned", key), c_init="uniform")
    stack = np.ascontiguousarray(w.stack[np.random.default_rng(5).permutation(w.stack_size)])
    out = run_stack(stream, w, stack=stack, device_transpose=False)
    assert oracle_error(w, out, stack=stack)[0] <= tol(dtype)


SHAPES_30K = [s for s in SHAPES if s != (23, 23, 23)]


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
@pytest.mark.parametrize("shape", SHAPES_30K, ids=["%dx%dx%d" % s for s in SHAPES_30K])
def test_full_size_shapes_30k_oracle(stream, shape, dtype):
    """The other workload shapes at the size the stack metric is quoted on: 30000 entries over 2000 A,
    2000 B and 1000 C blocks (the S >= 30000 launch geometry, no bs scaling), against the oracle."""
    m, n, k = shape
    w = make_stack_workload(m, n, k, 30000, nc=1000, na=2000, nb=2000, dtype=dtype, seed=_seed("30k", shape))
    out = run_stack(stream, w, device_transpose=False)
    err = oracle_error(w, out)
    assert err[0] <= tol(dtype), err


@pytest.mark.parametrize("case", ["one_run", "all_distinct", "single", "tail_run"])
def test_run_structure(stream, case):
    m, n, k = 13, 13, 13
    if case == "one_run":
        w = make_stack_workload(m, n, k, 5000, nc=1, dtype=np.float64, seed=11, c_init="uniform")
    elif case == "all_distinct":
        w = make_stack_workload(m, n, k, 2000, nc=2000, dtype=np.float64, seed=12, c_init="uniform", spread=0)
    elif case == "single":
        w = make_stack_workload(m, n, k, 1, nc=1, dtype=np.float64, seed=13, c_init="uniform")
    else:
        w = make_stack_workload(m, n, k, 3001, nc=10, dtype=np.float64, seed=14, c_init="uniform")
    out = run_stack(stream, w)
    assert oracle_error(w, out)[0] <= 1e-13


def _with_runs(w, lengths):
    """w's stack re-cut into C-runs of the given lengths (sorted by C, lengths sum to the stack size)."""
    assert sum(lengths) == w.stack_size and len(lengths) <= w.nc
    stack = w.stack.copy()
    stack[:, 2] = np.repeat(np.arange(len(lengths), dtype=np.int32), lengths) * (w.m * w.n) + 1
    return stack


LONG_RUNS = [1500, 3, 7, 40, 1, 64, 65, 33, 900] + [17] * 40 + [707]


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
@pytest.mark.parametrize("shape", [(23, 23, 23), (5, 5, 5), (13, 13, 13), (4, 9, 22)], ids=lambda s: "%dx%dx%d" % s)
def test_long_runs_cut(stream, shape, dtype):
    """Runs far longer than bs (one C block fed by many k-blocks, or the last C block of
    acc_bench's INIT_STACK taking the remainder) are cut into pieces at work-group boundaries and
    the pieces added atomically: every kernel variant, both flush modes, several bs."""
    m, n, k = shape
    tid = lx.typeid(dtype)
    S = sum(LONG_RUNS)
    w = make_stack_workload(m, n, k, S, nc=len(LONG_RUNS), dtype=dtype, seed=_seed("long", shape), c_init="uniform")
    stack = _with_runs(w, LONG_RUNS)
    checked = 0
    for variant, tm, tn, nt, vkt in lx.kernel_variants(tid):
        for bs in (5, 16, 64):
            for flush in (1, 2):
                cfg = lx.SmmConfig(variant=variant, tm=tm, tn=tn, nt=nt, bs=bs, kt=(vkt if variant else 0),
                                   flush=flush)
                if not lx.config_valid(tid, m, n, k, cfg):
                    continue
                out = run_stack(stream, w, config=cfg, device_transpose=False, stack=stack)
                err = oracle_error(w, out, stack=stack)[0]
                assert err <= tol(dtype), "v=%d tm=%d tn=%d nt=%d bs=%d flush=%d: %g" % (
                    variant, tm, tn, nt, bs, flush, err)
                checked += 1
    assert checked >= 6
    out = run_stack(stream, w, stack=stack)  # the tuned or default configuration
    assert oracle_error(w, out, stack=stack)[0] <= tol(dtype)


def test_long_k_slices(stream):
    """k beyond max_kernel_dim (allowed: the gate bounds m*n only) runs in k-slices."""
    m, n, k = 20, 24, 300
    w = make_stack_workload(m, n, k, 500, dtype=np.float64, seed=21, c_init="uniform")
    # k*n exceeds max_kernel_dim^2, so libsmm_acc_transpose leaves B alone (smm_trans.c:54-68):
    # upload B in the kernel's row-major layout directly
    out = run_stack(stream, w, device_transpose=False)
    assert oracle_error(w, out)[0] <= 1e-13


def test_repeat_accumulates(stream):
    """beta = 1 semantics: applying the stack twice accumulates twice."""
    m, n, k = 23, 23, 23
    w = make_stack_workload(m, n, k, 3000, dtype=np.float64, seed=22, c_init="uniform")
    out = run_stack(stream, w, repeat=2)
    assert oracle_error(w, out, repeat=2)[0] <= 1e-13


def test_full_size_23_30k_oracle(stream):
    """The 23^3 FP64 stack at full size: 2000 A, 2000 B, 1000 C blocks, 30000 entries."""
    w = make_stack_workload(23, 23, 23, 30000, nc=1000, na=2000, nb=2000, dtype=np.float64, seed=25071975)
    out = run_stack(stream, w)
    err = oracle_error(w, out)
    assert err[0] <= 1e-13, err


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
@pytest.mark.parametrize("shape", SHAPES, ids=["%dx%dx%d" % s for s in SHAPES])
def test_default_path_bitwise(stream, shape, dtype):
    """The shipped configuration, unmodified, on the 30000-entry stack of every workload shape (runs of
    26-34 entries): no run is cut, each C block is summed in one work-group in a fixed order and written
    once, so three calls give the same bits, and they match the oracle."""
    m, n, k = shape
    w = make_stack_workload(m, n, k, 30000, nc=1000, na=2000, nb=2000, dtype=dtype, seed=25071975)
    assert cut_limit(lx.typeid(dtype), m, n, k, w.stack_size) >= 34
    outs = [run_stack(stream, w, device_transpose=False) for _ in range(3)]
    bits = np.uint64 if dtype == np.float64 else np.uint32
    for out in outs[1:]:
        assert np.array_equal(outs[0].view(bits), out.view(bits))
    assert sampled_oracle_error(w, outs[0], nblocks=64)[0] <= tol(dtype)


@pytest.mark.parametrize("dtype", DTYPES, ids=["f64", "f32"])
@pytest.mark.parametrize("stack_size", [100000, 500000])
def test_full_size_checksum(stream, dtype, stack_size):
    """Large stacks: size-independent checks (checksum of the update; bitwise determinism for every
    C block whose run is not longer than split x bs of the launch, the others are cut into pieces added
    atomically)."""
    m = n = k = 23
    w = make_stack_workload(m, n, k, stack_size, dtype=dtype, seed=_seed("big", stack_size))
    out = run_stack(stream, w)
    total, scale = checksum(w)
    got = float(np.sum(out.astype(np.float64)))
    eps = 1e-12 if dtype == np.float64 else 1e-5
    assert abs(got - total) <= eps * scale, (got, total, scale)
    again = run_stack(stream, w)
    short = uncut_blocks(w, limit=cut_limit(lx.typeid(dtype), m, n, k, stack_size))
    assert np.array_equal(out[short], again[short])
    assert np.allclose(out[~short], again[~short], rtol=0, atol=eps * max(1.0, float(np.abs(out).max())) * 64)


LARGE = [(s, d, S) for s in SHAPES if s != (23, 23, 23) for d, S in ((np.float64, 100000), (np.float32, 100000),
                                                                       (np.float64, 500000))]


@pytest.mark.parametrize("shape,dtype,stack_size", LARGE,
                         ids=["%dx%dx%d-%s-%d" % (s + ("f64" if d == np.float64 else "f32", S)) for s, d, S in LARGE])
def test_full_size_checksum_shapes(stream, shape, dtype, stack_size):
    """The other workload shapes at 100k and 500k entries (pools S/30 C blocks, twice as many A and B
    blocks; up to 2.2 GB of 64^3 FP64 pools): the checksum of the update, and the double-double oracle on
    16 sampled C blocks (the entries that update them), through libsmm_acc_process with the shipped
    configuration."""
    m, n, k = shape
    w = make_stack_workload(m, n, k, stack_size, dtype=dtype, seed=_seed("large", shape, stack_size))
    out = run_stack(stream, w, device_transpose=False)
    total, scale = checksum(w)
    eps = 1e-12 if dtype == np.float64 else 1e-5
    assert abs(float(np.sum(out.astype(np.float64))) - total) <= eps * scale
    err = sampled_oracle_error(w, out, seed=stack_size)
    assert err[0] <= tol(dtype), err


def test_transpose_parity(stream):
    for dtype in DTYPES:
        for m, n in [(23, 23), (4, 9), (9, 4), (45, 13), (80, 80), (1, 7)]:
            rng = np.random.default_rng(m * 100 + n)
            nb = 300
            data = rng.uniform(-1, 1, size=nb * m * n).astype(dtype)
            trs = (rng.permutation(nb)[:200].astype(np.int32) * (m * n))
            offset = 17
            ref = oracle.transpose(m, n, trs, offset, data.copy())
            tid = lx.typeid(dtype)
            d = lx.DeviceBuffer.from_array(data)
            t = lx.DeviceBuffer.from_array(trs)
            lx.transpose(t, offset, len(trs) - offset, d, tid, m, n, stream)
            stream.sync()
            out = np.empty_like(data)
            d.download(out)
            d.free()
            t.free()
            assert np.array_equal(out.view(np.uint8), ref.view(np.uint8)), (dtype, m, n)


def test_transpose_round_trip(stream):
    """(k,n) then (n,k) is the identity (samples/smm/acc_bench.c:317-322)."""
    k, n, nb = 22, 9, 500
    data = np.random.default_rng(5).uniform(-1, 1, size=nb * k * n)
    d = lx.DeviceBuffer.from_array(data)
    t = lx.DeviceBuffer.from_array(np.arange(nb, dtype=np.int32) * (k * n))
    lx.transpose(t, 0, nb, d, lx.REAL_8, k, n, stream)
    lx.transpose(t, 0, nb, d, lx.REAL_8, n, k, stream)
    stream.sync()
    out = np.empty_like(data)
    d.download(out)
    d.free()
    t.free()
    assert np.array_equal(out, data)


def test_fallback_and_noop(stream):
    """-1 host fallback and successful no-ops (samples/smm/smm_kernel.c:524-526, smm_trans.c:54-68)."""
    w = make_stack_workload(23, 23, 23, 100, dtype=np.float64, seed=1)
    a = lx.DeviceBuffer.from_array(w.a)
    b = lx.DeviceBuffer.from_array(w.b)
    c = lx.DeviceBuffer.from_array(w.c)
    st = lx.DeviceBuffer.from_array(w.stack)
    try:
        assert lx.process(st, 100, lx.REAL_8, a, b, c, 23, 23, 23, stream, def_mnk=0) == -1
        assert lx.process(st, 100, lx.REAL_8, a, b, c, 81, 81, 23, stream) == -1
        assert lx.process(st, 100, lx.COMPLEX_8, a, b, c, 23, 23, 23, stream) == -1
        assert lx.process(st, 0, lx.REAL_8, a, b, c, 23, 23, 23, stream) == 0
        before = np.empty_like(w.c)
        c.download(before)
        assert lx.transpose(st, 0, 0, b, lx.REAL_8, 23, 23, stream) == 0
        assert lx.transpose(st, 0, 10, b, lx.COMPLEX_8, 23, 23, stream) == 0
        stream.sync()
        after = np.empty_like(w.c)
        c.download(after)
        assert np.array_equal(before, after)
    finally:
        for buf in (a, b, c, st):
            buf.free()


def test_memory_api(stream):
    x = np.arange(1000, dtype=np.float64)
    d = lx.DeviceBuffer.from_array(x)
    e = lx.DeviceBuffer(x.nbytes)
    lx.check(lx.lib().libxstream_mem_copy_d2d(d.ptr, e.ptr, x.nbytes, stream.handle))
    e.zero(stream, offset=80, nbytes=160)  # elements 10..29
    stream.sync()
    out = np.empty_like(x)
    e.download(out)
    ref = x.copy()
    ref[10:30] = 0
    assert np.array_equal(out, ref)
    # odd byte offset and length exercise the unaligned edges of the zero fill
    e.upload(x)
    e.zero(None, offset=3, nbytes=4001)
    raw = np.empty(x.nbytes, dtype=np.uint8)
    e.download(raw)
    ref_raw = x.view(np.uint8).copy()
    ref_raw[3:4004] = 0
    assert np.array_equal(raw, ref_raw)
    d.free()
    e.free()
    ev0, ev1 = lx.Event(), lx.Event()
    ev0.record(stream)
    ev1.record(stream)
    ev1.sync()
    assert ev0.elapsed_ms(ev1) >= 0
    ev0.close()
    ev1.close()
