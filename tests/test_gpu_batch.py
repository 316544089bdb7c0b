"""GPU parity of the batched mixed-shape launch (one persistent kernel), eager and hipGraph replays.

Every stack of the batch is checked against the oracle with the libsmm_acc_process bar (max
normalised error <= 1e-13 FP64 / 1e-5 FP32). Graph replays are checked one by one, with the
program's own eager work (downloads, synchronisations) between them.
"""
import numpy as np
import pytest

import libxstream_amd as lx
import oracle
from libxstream_amd.workload import make_mixed_batch
from tests.helpers import tol

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    s = lx.Stream("batch")
    yield s
    s.close()
    lx.finalize()


class Resident:
    def __init__(self, works, stream):
        self.works = works
        self.bufs = []
        self.items = []
        trs_cache = {}
        for w in works:
            tid = lx.typeid(w.dtype)
            a = lx.DeviceBuffer.from_array(w.a)
            b = lx.DeviceBuffer.from_array(w.b)
            trs = lx.DeviceBuffer.from_array(np.arange(w.nb, dtype=np.int32) * (w.k * w.n))
            lx.transpose(trs, 0, w.nb, b, tid, w.k, w.n, stream)
            c = lx.DeviceBuffer.from_array(w.c)
            st = lx.DeviceBuffer.from_array(w.stack)
            self.bufs += [a, b, trs, c, st]
            self.items.append((st, w.stack_size, w.m, w.n, w.k, a, b, c))
        stream.sync()

    def download(self, i):
        out = np.empty_like(self.works[i].c)
        self.items[i][7].download(out)
        return out

    def free(self):
        for b in self.bufs:
            b.free()


def check_all(res, expected):
    for i, w in enumerate(res.works):
        out = res.download(i)
        hi, lo = expected[i]
        err = oracle.error(out.ravel(), hi, lo)
        assert err[0] <= tol(w.dtype), "stack %d (%dx%dx%d, S=%d): %g" % (i, w.m, w.n, w.k, w.stack_size, err[0])


@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_batch_eager(stream, dtype):
    works = make_mixed_batch(24, seed=11, dtype=dtype, len_range=(20, 3000), c_init="uniform")
    res = Resident(works, stream)
    try:
        batch = lx.Batch(lx.typeid(dtype), res.items)
        info = batch.info()
        assert info["ntasks"] > 0 and info["flops"] == sum(w.flops() for w in works)
        batch.process(stream)
        stream.sync()
        expected = [oracle.stack_dd(w.m, w.n, w.k, w.stack, w.a, w.b, c_in=w.c.ravel()) for w in works]
        check_all(res, expected)
        batch.close()
    finally:
        res.free()


def test_batch_graph_replays(stream):
    """Three replays of the captured launch, each checked, eager work in between."""
    works = make_mixed_batch(24, seed=12, dtype=np.float64, len_range=(20, 2000), c_init="uniform")
    res = Resident(works, stream)
    try:
        batch = lx.Batch(lx.REAL_8, res.items)
        c_ref = [w.c.ravel().copy() for w in works]
        for replay in range(3):
            batch.replay(stream)
            stream.sync()
            expected = []
            for i, w in enumerate(works):
                hi, lo = oracle.stack_dd(w.m, w.n, w.k, w.stack, w.a, w.b, c_in=c_ref[i])
                expected.append((hi, lo))
                c_ref[i] = res.download(i).ravel().copy()  # next replay starts from the device state
            check_all(res, expected)
        batch.close()
    finally:
        res.free()


def test_batch_matches_process(stream):
    """Shapes with more than 4 output tiles (team of wavefronts) and tiny ones, unsorted stacks too."""
    from libxstream_amd.workload import make_stack_workload

    works = [make_stack_workload(m, n, k, s, dtype=np.float64, seed=5 + m, c_init="uniform")
             for (m, n, k, s) in [(64, 64, 64, 900), (50, 33, 7, 1500), (4, 4, 4, 3000), (17, 64, 40, 800),
                                  (1, 1, 1, 500), (64, 3, 64, 700)]]
    works[2].stack[:] = works[2].stack[np.random.default_rng(1).permutation(works[2].stack_size)]
    res = Resident(works, stream)
    try:
        batch = lx.Batch(lx.REAL_8, res.items)
        batch.process(stream)
        stream.sync()
        expected = [oracle.stack_dd(w.m, w.n, w.k, w.stack, w.a, w.b, c_in=w.c.ravel()) for w in works]
        check_all(res, expected)
        batch.close()
    finally:
        res.free()


def test_batch_long_runs(stream):
    """Runs much longer than a chunk are cut at chunk boundaries; pieces add atomically."""
    from libxstream_amd.workload import make_stack_workload

    lengths = [[2500, 3, 700, 40], [1, 1200, 64, 65, 1], [3000]]
    works = []
    for j, (m, n, k) in enumerate([(23, 23, 23), (64, 64, 64), (5, 7, 9)]):
        w = make_stack_workload(m, n, k, sum(lengths[j]), nc=len(lengths[j]), dtype=np.float64, seed=31 + j,
                                c_init="uniform")
        w.stack[:, 2] = np.repeat(np.arange(len(lengths[j]), dtype=np.int32), lengths[j]) * (m * n) + 1
        works.append(w)
    res = Resident(works, stream)
    try:
        batch = lx.Batch(lx.REAL_8, res.items, lx.BatchParams(chunk_elements=8 * 2 * 23 * 23, blocks_per_cu=8))
        batch.process(stream)
        stream.sync()
        expected = [oracle.stack_dd(w.m, w.n, w.k, w.stack, w.a, w.b, c_in=w.c.ravel()) for w in works]
        check_all(res, expected)
        batch.close()
    finally:
        res.free()


@pytest.mark.parametrize("mode", [1, 2], ids=["in_order", "lanes"])
def test_batch_per_stack_mode(stream, mode):
    """per_stack = 1: one libsmm_acc_process launch per stack, in order; 2: the same launches spread over
    side streams that fork from the caller's stream and join it (concurrent branches of one graph). Eager,
    then three captured-graph replays, each checked with eager work in between."""
    works = make_mixed_batch(12, seed=17, dtype=np.float64, len_range=(20, 3000), c_init="uniform")
    res = Resident(works, stream)
    try:
        batch = lx.Batch(lx.REAL_8, res.items, lx.BatchParams(per_stack=mode))
        c_ref = [w.c.ravel().copy() for w in works]
        for step in range(4):
            if step == 0:
                batch.process(stream)
            else:
                batch.replay(stream)
            stream.sync()
            expected = []
            for i, w in enumerate(works):
                expected.append(oracle.stack_dd(w.m, w.n, w.k, w.stack, w.a, w.b, c_in=c_ref[i]))
                c_ref[i] = res.download(i).ravel().copy()
            check_all(res, expected)
        batch.close()
    finally:
        res.free()


@pytest.mark.parametrize("per_stack", [0, 1, 2])
def test_batch_unsorted_stacks(stream, per_stack):
    """Stacks whose C-runs are interleaved (legal input): both batch modes flush with atomic adds, so
    every stack still matches the oracle, whatever configuration the per-stack mode resolves."""
    works = make_mixed_batch(10, seed=29, dtype=np.float64, len_range=(200, 3000), c_init="uniform")
    rng = np.random.default_rng(31)
    for w in works:
        w.stack = np.ascontiguousarray(w.stack[rng.permutation(w.stack_size)])
    works.append(make_mixed_batch(1, seed=37, dtype=np.float64, dim_range=(45, 45), len_range=(3000, 3000),
                                  c_init="uniform")[0])  # 45x45x45: a wide block, runs interleaved below
    works[-1].stack = np.ascontiguousarray(works[-1].stack[rng.permutation(works[-1].stack_size)])
    res = Resident(works, stream)
    try:
        expected = [oracle.stack_dd(w.m, w.n, w.k, w.stack, w.a, w.b, c_in=w.c.ravel()) for w in works]
        batch = lx.Batch(lx.REAL_8, res.items, lx.BatchParams(per_stack=per_stack))
        batch.process(stream)
        stream.sync()
        check_all(res, expected)
        batch.close()
    finally:
        res.free()


def test_batch_rejects_large_blocks(stream):
    w = np.zeros(8)
    d = lx.DeviceBuffer.from_array(w)
    try:
        with pytest.raises(lx.LibxstreamError):
            lx.Batch(lx.REAL_8, [(d, 1, 65, 8, 8, d, d, d)])
    finally:
        d.free()


def test_batch_tune_schedule(stream):
    """The schedule autotuner leaves the caller's C alone, and every schedule gives the oracle's result."""
    works = make_mixed_batch(12, seed=13, dtype=np.float64, len_range=(50, 4000), c_init="uniform")
    res = Resident(works, stream)
    try:
        before = [res.download(i).copy() for i in range(len(works))]
        best, ms = lx.Batch.tune(lx.REAL_8, res.items)
        # either schedule may win: per-stack launches, or the persistent kernel with its chunk size and
        # work-groups per CU
        assert ms > 0 and (best.per_stack in (1, 2) or (best.chunk_elements > 0 and best.blocks_per_cu > 0))
        for i in range(len(works)):
            assert np.array_equal(res.download(i), before[i]), "tuning wrote the caller's C (stack %d)" % i
        expected = [oracle.stack_dd(w.m, w.n, w.k, w.stack, w.a, w.b, c_in=w.c.ravel()) for w in works]
        for params in (best, lx.BatchParams(8 * 2 * 23 * 23, 4), lx.BatchParams(64 * 2 * 23 * 23, 16)):
            for i, w in enumerate(works):
                res.items[i][7].upload(w.c)
            batch = lx.Batch(lx.REAL_8, res.items, params)
            batch.replay(stream)
            stream.sync()
            check_all(res, expected)
            batch.close()
    finally:
        res.free()


def test_batch_full_workload(stream):
    """The full mixed-shape workload (24 stacks, seed 25071975, lengths 1k-200k: about 915k entries), in
    the schedule libsmm_acc_batch_tune picks and in one persistent launch: eager, then two graph replays.
    Each stack's checksum after every application, and the double-double oracle on 8 sampled C blocks per
    stack after the first."""
    from tests.helpers import checksum, sampled_oracle_error

    works = make_mixed_batch(24, seed=25071975, dtype=np.float64)
    assert sum(w.stack_size for w in works) > 500000
    totals = [checksum(w) for w in works]
    res = Resident(works, stream)
    try:
        best, _ = lx.Batch.tune(lx.REAL_8, res.items)
        for params in (best, lx.BatchParams(0, 0, 0), lx.BatchParams(0, 0, 2)):
            for i, w in enumerate(works):
                res.items[i][7].upload(w.c)
            batch = lx.Batch(lx.REAL_8, res.items, params)
            for step in range(3):
                if step == 0:
                    batch.process(stream)
                else:
                    batch.replay(stream)
                stream.sync()
                for i, w in enumerate(works):
                    out = res.download(i)
                    total, scale = totals[i]
                    got = float(np.sum(out))
                    assert abs(got - (step + 1) * total) <= 1e-12 * (step + 1) * scale, (params.as_dict(), step, i)
                    if step == 0:
                        err = sampled_oracle_error(w, out, nblocks=8, seed=i)
                        assert err[0] <= tol(w.dtype), (params.as_dict(), i, err)
            batch.close()
    finally:
        res.free()
