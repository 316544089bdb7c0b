"""Concurrent use from several host threads (libsmm_acc_is_thread_safe returns 1, as the reference's
smm_acc.c:484-491 does under OpenMP): the configuration registry and its lookups on the CPU, and
concurrent stack launches on per-thread streams on the GPU."""
import threading

import numpy as np
import pytest

import libxstream_amd as lx
from libxstream_amd.workload import make_stack_workload
from tests.helpers import oracle_error, tol


def _run_threads(fn, nthreads):
    errors = []

    def body(t):
        try:
            fn(t)
        except Exception as e:  # reported in the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=body, args=(t,)) for t in range(nthreads)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors


def test_registry_concurrent_set_get():
    """Eight threads register and read back distinct shapes while all of them resolve shared ones."""
    def work(t):
        for i in range(200):
            m, n, k = 4 + t, 5 + i % 50, 6 + i % 7
            cfg = lx.SmmConfig(variant=0, tm=1, tn=1, nt=64, bs=8 + t, kt=0, flush=1, s=1000 + i)
            if lx.config_valid(lx.REAL_8, m, n, k, cfg):
                lx.config_set(lx.REAL_8, m, n, k, cfg)
                got, src = lx.config_get(lx.REAL_8, m, n, k)
                assert src == 1 and (got.bs, got.s) == (8 + t, 1000 + i), (t, i, got)
            shared, _ = lx.config_get(lx.REAL_8, 23, 23, 23)
            assert lx.config_valid(lx.REAL_8, 23, 23, 23, shared)

    _run_threads(work, 8)


@pytest.mark.gpu
def test_concurrent_process_streams():
    """Four threads, each with its own stream and shapes (first calls resolve configurations concurrently)."""
    lx.init()
    assert lx.device_count() > 0
    shapes = [(23, 23, 23), (13, 13, 13), (4, 9, 22), (45, 45, 13), (5, 5, 5), (32, 32, 32), (31, 17, 9),
              (7, 61, 40)]
    works = [make_stack_workload(m, n, k, 2500, dtype=np.float64, seed=100 + i, c_init="uniform")
             for i, (m, n, k) in enumerate(shapes)]
    outs = [None] * len(works)

    def work(t):
        stream = lx.Stream("t%d" % t)
        bufs = []
        try:
            for i in range(t, len(works), 4):
                w = works[i]
                a, b, c = (lx.DeviceBuffer.from_array(v) for v in (w.a, w.b_device_layout(), w.c))
                st = lx.DeviceBuffer.from_array(w.stack)
                bufs += [a, b, c, st]
                for _ in range(3):  # the stack applied three times on this thread's stream
                    assert lx.process(st, w.stack_size, lx.REAL_8, a, b, c, w.m, w.n, w.k, stream) == 0
                stream.sync()
                out = np.empty_like(w.c)
                c.download(out)
                outs[i] = out
        finally:
            for x in bufs:
                x.free()
            stream.close()

    _run_threads(work, 4)
    for i, w in enumerate(works):
        assert oracle_error(w, outs[i], repeat=3)[0] <= tol(np.float64), shapes[i]
