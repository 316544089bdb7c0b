#!/usr/bin/env python3
"""Benchmark: DBCSR stack multiply (libsmm_acc_process) on one MI355X, FP64 23x23x23.

One step = one libsmm_acc_process call over the whole parameter stack (inputs resident in HBM),
like one repetition of samples/smm/acc_bench.c:338-342. Prints one JSON line.

  value         whole-job GFLOP/s = 2*m*n*k*S / wall time per step (host clock, stream-synchronised)
  roofline      the stack kernel's algorithmic bytes S*8*(mk+kn+2mn) (samples/smm/smm_kernel.c:504-506)
                per launch / its average duration from HIP events recorded on the launch stream,
                against 8 TB/s HBM; `traffic` is the HBM bytes per launch read by rocprofv3 PMC
                counters (FETCH_SIZE, WRITE_SIZE) in two child runs, each scaled by what the same
                counter reads for a device-to-device copy of known size in the same run
  cpu_baseline  the reference host path (OpenMP, whole C-runs per thread) on the host's cores,
                timed on a bounded sample of the same stack
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
COPY_BYTES = 256 << 20  # calibration copy for the PMC byte counters


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--m", type=int, default=23)
    p.add_argument("--n", type=int, default=None)
    p.add_argument("--k", type=int, default=None)
    p.add_argument("--stack-size", type=int, default=500000)
    p.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    p.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU-baseline time (0 = skip)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 PMC child runs")
    p.add_argument("--no-small", action="store_true", help="skip the 30000-entry stack line")
    p.add_argument("--no-batch", action="store_true", help="skip the batched mixed-shape measurement")
    p.add_argument("--no-ozaki", action="store_true", help="skip the Ozaki-int8 DGEMM measurement")
    p.add_argument("--ozaki-sizes", default="1024,4096,8192", help="square DGEMM sizes for the Ozaki section")
    p.add_argument("--ozaki-only", action="store_true", help="print only the Ozaki section (no stack measurement)")
    p.add_argument("--no-learned", action="store_true", help="skip the learned-predictor evaluation")
    p.add_argument("--sweep", action="store_true",
                   help="print only the plain-stack sweep: 7 shapes x FP64/FP32 x S in {30k, 100k, 500k}")
    p.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    return p.parse_args()


def make_workload(args):
    from libxstream_amd.workload import make_stack_workload

    m = args.m
    n = args.n or m
    k = args.k or m
    dtype = np.float64 if args.dtype == "f64" else np.float32
    return make_stack_workload(m, n, k, args.stack_size, dtype=dtype, seed=25071975)


class Resident:
    """The workload's pools and stack uploaded once; B transposed on the device (acc_bench.c:325)."""

    def __init__(self, lx, w, stream):
        self.lx, self.w = lx, w
        self.tid = lx.typeid(w.dtype)
        self.a = lx.DeviceBuffer.from_array(w.a)
        self.b = lx.DeviceBuffer.from_array(w.b)
        self.trs = lx.DeviceBuffer.from_array(np.arange(w.nb, dtype=np.int32) * (w.k * w.n))
        self.c = lx.DeviceBuffer.from_array(w.c)
        self.st = lx.DeviceBuffer.from_array(w.stack)
        lx.transpose(self.trs, 0, w.nb, self.b, self.tid, w.k, w.n, stream)

    def process(self, stream):
        w = self.w
        rc = self.lx.process(self.st, w.stack_size, self.tid, self.a, self.b, self.c, w.m, w.n, w.k, stream)
        assert rc == 0, "libsmm_acc_process returned %d" % rc

    def free(self):
        for buf in (self.a, self.b, self.trs, self.c, self.st):
            buf.free()


def timed(lx, res, stream, steps, warmup):
    for _ in range(warmup):
        res.process(stream)
    stream.sync()
    lx.profile_enable(True)
    lx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        res.process(stream)
    stream.sync()
    t1 = time.perf_counter()
    prof = lx.profile_get("smm")
    lx.profile_enable(False)
    return (t1 - t0) * 1e3 / steps, prof["ms"] / max(prof["count"], 1)


def copyin_section(lx, w, stream, reps=5):
    """Copy-in of the A and B pools and the stack from pinned host memory (acc_bench.c:288-307): the
    three host-to-device copies enqueued on one stream, timed over reps after one untimed round."""
    names = ("a", "b", "stack")
    arrays = (w.a, w.b, w.stack)
    host = [lx.HostBuffer(x.dtype, x.size, stream) for x in arrays]
    dev = [lx.DeviceBuffer(x.nbytes) for x in arrays]
    try:
        for h, x in zip(host, arrays):
            h.array[...] = x.ravel()
        L = lx.lib()

        def copy_all():
            for h, d, x in zip(host, dev, arrays):
                check = L.libxstream_mem_copy_h2d(h.ptr, d.ptr, x.nbytes, stream.handle)
                assert check == 0, "libxstream_mem_copy_h2d returned %d" % check

        copy_all()
        stream.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            copy_all()
        stream.sync()
        sec = (time.perf_counter() - t0) / reps
    finally:
        for buf in host + dev:
            buf.free()
    size = float(sum(x.nbytes for x in arrays))
    return {"arrays": list(names), "bytes": size, "ms": round(sec * 1e3, 4), "gbs": round(size / sec * 1e-9, 2),
            "gibs": round(size / (sec * (1 << 30)), 2), "pinned": True}


def transfer_section(lx, m, n, k, dtype, stream, nstacks=16, rounds=3):
    """Per-call stack copy (DBCSR's pattern, libxstream_mem.c:1094-1127): nstacks different 30000-entry
    stacks over resident 2000/2000/1000-block pools, each copied from pinned staging memory and then
    multiplied. Serialised: one stream, one slot (copy, multiply, copy, ...). Overlapped: the copies on
    their own stream through 2 slots (and 4, for comparison), so the copy of stack i+1 runs during the
    multiply of stack i."""
    from libxstream_amd.transfer import StackPipeline
    from libxstream_amd.workload import make_stack_workload

    w = make_stack_workload(m, n, k, 30000, nc=1000, na=2000, nb=2000, dtype=dtype, seed=25071975)
    stacks = [make_stack_workload(m, n, k, 30000, nc=1000, na=2000, nb=2000, dtype=dtype, seed=25071975 + i).stack
              for i in range(nstacks)]
    res = Resident(lx, w, stream)
    upload = lx.Stream("upload")
    out = {"stacks": nstacks, "entries": 30000, "stack_bytes": int(stacks[0].nbytes)}
    try:
        def launch(dev, entries):
            rc = lx.process(dev, entries, res.tid, res.a, res.b, res.c, m, n, k, stream)
            assert rc == 0, "libsmm_acc_process returned %d" % rc

        for name, pipe in (("serialised", StackPipeline(30000, stream, None, nslots=1)),
                           ("overlapped", StackPipeline(30000, stream, upload, nslots=2)),
                           ("overlapped_4slots", StackPipeline(30000, stream, upload, nslots=4))):
            try:
                for st in stacks[:2]:  # warm-up
                    pipe.submit(st, launch)
                pipe.sync()
                t0 = time.perf_counter()
                for _ in range(rounds):
                    for st in stacks:
                        pipe.submit(st, launch)
                pipe.sync()
                out[name + "_ms_per_stack"] = round((time.perf_counter() - t0) * 1e3 / (rounds * nstacks), 4)
            finally:
                pipe.close()
        out["kernel_gflops_overlapped"] = round(w.flops() / (out["overlapped_ms_per_stack"] * 1e-3) * 1e-9, 1)
        out["overlap_gain"] = round(out["serialised_ms_per_stack"] / out["overlapped_ms_per_stack"], 3)
    finally:
        upload.close()
        res.free()
    return out


def transpose_section(lx, res, stream, pairs=10):
    """libsmm_acc_transpose over the whole B pool (smm_trans.c:48-228): kernel time from the launch records
    and HBM rate over the reference's bytes, 2 S m n ts per launch (smm_trans.c:208-212). Transposes go
    in pairs, (k, n) then (n, k), so B is left in the kernel layout."""
    w = res.w
    for _ in range(2):
        lx.transpose(res.trs, 0, w.nb, res.b, res.tid, w.k, w.n, stream)
        lx.transpose(res.trs, 0, w.nb, res.b, res.tid, w.n, w.k, stream)
    stream.sync()
    lx.profile_enable(True)
    lx.profile_reset()
    for _ in range(pairs):
        lx.transpose(res.trs, 0, w.nb, res.b, res.tid, w.k, w.n, stream)
        lx.transpose(res.trs, 0, w.nb, res.b, res.tid, w.n, w.k, stream)
    stream.sync()
    prof = lx.profile_get("trans")
    lx.profile_enable(False)
    ms = prof["ms"] / max(prof["count"], 1)
    nbytes = 2.0 * w.nb * w.k * w.n * w.a.itemsize
    return {"blocks": w.nb, "shape": [w.k, w.n], "kernel_ms": round(ms, 5), "bytes_per_launch": nbytes,
            "gbs": round(nbytes / (ms * 1e-3) * 1e-9, 1), "frac_of_8tbs": round(nbytes / (ms * 1e-3) * 1e-9 / HBM_PEAK_GBS, 4)}


SWEEP_SHAPES = ((5, 5, 5), (13, 13, 13), (23, 23, 23), (32, 32, 32), (64, 64, 64), (4, 9, 22), (45, 45, 13))


def sweep_section(lx, stream, steps=10, warmup=2):
    """Plain stacks over the sweep shapes, FP64 and FP32, S = 30k, 100k, 500k (the bench's generator: pools
    proportional to S, nc = S/30, na = nb = 2 nc, runs of 30 +- 4): kernel time from the launch records,
    FLOP/s and algorithmic GB/s (S ts (mk + kn + 2mn), samples/smm/smm_kernel.c:504-506) against 8 TB/s."""
    from libxstream_amd.workload import make_stack_workload

    rows = []
    for dtype in (np.float64, np.float32):
        for (m, n, k) in SWEEP_SHAPES:
            for S in (30000, 100000, 500000):
                w = make_stack_workload(m, n, k, S, dtype=dtype, seed=25071975)
                res = Resident(lx, w, stream)
                _, kernel_ms = timed(lx, res, stream, steps, warmup)
                cfg, src_ = lx.config_get(lx.typeid(dtype), m, n, k)
                res.free()
                gbs = w.bytes_alg() / (kernel_ms * 1e-3) * 1e-9
                rows.append({"dtype": "f64" if dtype == np.float64 else "f32", "m": m, "n": n, "k": k,
                             "stack_size": S, "kernel_ms": round(kernel_ms, 5),
                             "gflops": round(w.flops() / (kernel_ms * 1e-3) * 1e-9, 1),
                             "alg_gbs": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                             "pool_mb": round((w.a.nbytes + w.b.nbytes + w.c.nbytes) / 1e6, 1),
                             "variant": cfg.as_dict().get("variant"), "config_source":
                             ["default", "tuned", "env", "learned"][src_]})
                print("sweep %s %dx%dx%d S=%d: %.4f ms, %.1f GB/s" % (rows[-1]["dtype"], m, n, k, S, kernel_ms, gbs),
                      file=sys.stderr, flush=True)
                del w, res
    return rows


def batched(lx, stream, steps, warmup):
    """24 stacks, (m, n, k) and lengths log-uniform in 4..64 and 1k..200k, FP64: one persistent launch
    replayed from a hipGraph, against the same stacks as 24 libsmm_acc_process launches."""
    from libxstream_amd.workload import make_mixed_batch

    works = make_mixed_batch(24, seed=25071975, dtype=np.float64)
    res = [Resident(lx, w, stream) for w in works]
    stream.sync()
    items = [(r.st, r.w.stack_size, r.w.m, r.w.n, r.w.k, r.a, r.b, r.c) for r in res]
    # the default schedule, then the autotuned one (libsmm_acc_batch_tune: timed on these stacks, C untouched)
    batch = lx.Batch(lx.REAL_8, items)
    for _ in range(warmup):
        batch.replay(stream)
    stream.sync()
    lx.profile_enable(True)
    lx.profile_reset()
    for _ in range(steps):
        batch.replay(stream)
    stream.sync()
    prof0 = lx.profile_get("batch")
    lx.profile_enable(False)
    batch.close()
    best, _ = lx.Batch.tune(lx.REAL_8, items)
    batch = lx.Batch(lx.REAL_8, items, best)
    info = batch.info()
    for _ in range(warmup):
        batch.replay(stream)
    stream.sync()
    lx.profile_enable(True)
    lx.profile_reset()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.replay(stream)
    stream.sync()
    t1 = time.perf_counter()
    prof = lx.profile_get("batch")
    for _ in range(warmup):
        for r in res:
            r.process(stream)
    stream.sync()
    t2 = time.perf_counter()
    for _ in range(steps):
        for r in res:
            r.process(stream)
    stream.sync()
    t3 = time.perf_counter()
    per_stack = []  # kernel time of each stack's own launch: where the batch's time goes
    for r in res:
        lx.profile_reset()
        for _ in range(steps):
            r.process(stream)
        stream.sync()
        p = lx.profile_get("smm")
        ms = p["ms"] / max(p["count"], 1)
        per_stack.append({"shape": "%dx%dx%d" % (r.w.m, r.w.n, r.w.k), "entries": r.w.stack_size,
                          "kernel_ms": round(ms, 4), "hbm_frac": round(r.w.bytes_alg() / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 3)})
    lx.profile_enable(False)
    batch.close()
    for r in res:
        r.free()
    graph_ms = (t1 - t0) * 1e3 / steps
    single_ms = (t3 - t2) * 1e3 / steps
    return {
        "workload": "24 FP64 stacks, (m,n,k) log-uniform in 4..64, lengths log-uniform 1k..200k, seed 25071975",
        "entries": int(sum(w.stack_size for w in works)),
        "shapes": ["%dx%dx%d:%d" % (w.m, w.n, w.k, w.stack_size) for w in works],
        "tasks": info["ntasks"],
        "graph_ms_per_step": round(graph_ms, 4),
        "graph_gflops": round(info["flops"] / (graph_ms * 1e-3) * 1e-9, 1),
        "graph_kernel_ms": round(prof["ms"] / max(prof["count"], 1), 4),
        "schedule": dict(best.as_dict(), source="libsmm_acc_batch_tune"),
        "default_schedule_graph_kernel_ms": round(prof0["ms"] / max(prof0["count"], 1), 4),
        "alg_gbs": round(info["bytes"] / (prof["ms"] / max(prof["count"], 1) * 1e-3) * 1e-9, 1),
        "separate_launches_ms_per_step": round(single_ms, 4),
        "separate_launches_gflops": round(info["flops"] / (single_ms * 1e-3) * 1e-9, 1),
        "separate_kernels": per_stack,
    }


def _rowcol(rng, rows, cols, e, per_row):
    """mantissas uniform +-[1,2) times 2^x, x uniform in [-e, e] per row (per_row) or per column"""
    mant = rng.uniform(1.0, 2.0, (rows, cols))
    mant *= np.where(rng.integers(0, 2, (rows, cols), dtype=np.int8) != 0, 1.0, -1.0)
    return mant * np.exp2(rng.integers(-e, e + 1, (rows, 1) if per_row else (1, cols))).astype(np.float64)


def _sampled_error(A, B, c_colmajor, rng, count=2048):
    """normwise error over sampled entries against 80-bit extended-precision dot products"""
    if np.finfo(np.longdouble).nmant < 63:
        return None
    m, n = A.shape[0], B.shape[1]
    i, j = rng.integers(0, m, count), rng.integers(0, n, count)
    ref = (A[i].astype(np.longdouble) * B[:, j].T.astype(np.longdouble)).sum(axis=1)
    got = c_colmajor[i + j * m].astype(np.longdouble)
    return float(np.sqrt(((got - ref) ** 2).sum() / (ref ** 2).sum()))


class _RocBlas:
    """The vendor FP64 GEMM (rocBLAS) as a second comparator; None when the library is absent."""

    def __init__(self, hip_stream):
        import ctypes

        self.ct = ctypes
        self.lib = None
        for name in ("librocblas.so", "/opt/rocm/lib/librocblas.so"):
            try:
                self.lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.lib is None:
            return
        self.handle = ctypes.c_void_p()
        assert self.lib.rocblas_create_handle(ctypes.byref(self.handle)) == 0
        assert self.lib.rocblas_set_stream(self.handle, ctypes.c_void_p(hip_stream)) == 0
        self.lib.rocblas_dgemm.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p,
                                           ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double), ctypes.c_void_p, ctypes.c_int]

    def dgemm(self, n, a, b, c):
        one, zero = self.ct.c_double(1.0), self.ct.c_double(0.0)
        rc = self.lib.rocblas_dgemm(self.handle, 111, 111, n, n, n, self.ct.byref(one), a.ptr, n, b.ptr, n,
                                    self.ct.byref(zero), c.ptr, n)
        assert rc == 0, "rocblas_dgemm returned %d" % rc

    def close(self):
        if self.lib is not None:
            self.lib.rocblas_destroy_handle(self.handle)


def _event_ms(lx, stream, fn, reps):
    fn()  # warmup (first call grows the scratch)
    stream.sync()
    e0, e1 = lx.Event(), lx.Event()
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.sync()
    ms = e0.elapsed_ms(e1) / reps
    e0.close()
    e1.close()
    return ms


# (kind, slices or primes, trim): scheme II with up to 16 moduli (the reference's default kind 3 picks it at
# these sizes) on the full 53-bit mantissas and with 2 and 6 bits trimmed (one and two moduli fewer), scheme I
# over the slice triangle (cutoff s - 1) for s = 3 .. 8, and scheme I at the reference's default 8 slices over
# all pairs
OZAKI_CONFIGS = ((2, 16, 0), (2, 16, 1), (2, 16, 3)) + tuple((1, s, s - 1) for s in range(3, 9)) + ((1, 8, 0),)


def ozaki_section(lx, stream, sizes, reps=3):
    """Square DGEMM N x N x N: native FP64 matrix cores (own kernel and rocBLAS) against Ozaki-int8."""
    rng = np.random.default_rng(25071975)
    roc = _RocBlas(stream.native())
    results = []
    for n in sizes:
        flops = 2.0 * n ** 3
        entry = {"n": n}
        for e in ((10, 0, 30) if n <= 4096 else (10,)):  # e = 10 first: it carries the timing
            A, B = _rowcol(rng, n, n, e, True), _rowcol(rng, n, n, e, False)
            a, b = np.asfortranarray(A).ravel(order="F"), np.asfortranarray(B).ravel(order="F")
            da, db = lx.DeviceBuffer.from_array(a), lx.DeviceBuffer.from_array(b)
            dc = lx.DeviceBuffer(8 * n * n)
            out = np.empty(n * n)
            if e == 10:  # timing (independent of the exponent range)
                ms = _event_ms(lx, stream, lambda: lx.dgemm_native("N", "N", n, n, n, 1.0, da, n, db, n, 0.0, dc, n,
                                                                   stream), reps)
                dc.download(out)
                entry["native_ms"] = round(ms, 4)
                entry["native_tflops"] = round(flops / ms * 1e-9, 2)
                entry["native_error"] = _sampled_error(A, B, out, rng)
                if roc.lib is not None:
                    ms = _event_ms(lx, stream, lambda: roc.dgemm(n, da, db, dc), reps)
                    dc.download(out)
                    entry["rocblas_ms"] = round(ms, 4)
                    entry["rocblas_tflops"] = round(flops / ms * 1e-9, 2)
                    entry["rocblas_error"] = _sampled_error(A, B, out, rng)
                entry["configs"] = []
            for cfg_i, (kind, sl, trim) in enumerate(OZAKI_CONFIGS):
                oz = lx.Ozaki(nslices=sl, trim=trim) if kind == 1 else lx.Ozaki(kind=kind, nprimes=sl, trim=trim)
                pairs = sum(1 for p in range(sl) for q in range(sl) if p + q <= 2 * (sl - 1) - trim)
                if e == 10:
                    lx.profile_enable(True)
                    lx.profile_reset()
                    ms = _event_ms(lx, stream, lambda: oz.gemm("N", "N", n, n, n, 1.0, da, n, db, n, 0.0, dc, n,
                                                               stream), reps)
                    i8 = lx.profile_get("ozaki_i8")
                    pre = lx.profile_get("ozaki_pre")
                    rec = lx.profile_get("ozaki_fp64")
                    lx.profile_enable(False)
                    i8_ms = i8["ms"] / max(i8["count"], 1)
                    best_native = min(entry["native_ms"], entry.get("rocblas_ms", float("inf")))
                    products = pairs if kind == 1 else int(round(i8["flops"] / max(i8["count"], 1) / (2.0 * n ** 3)))
                    entry["configs"].append({
                        "scheme": kind, ("slices" if kind == 1 else "max_primes"): sl, "trim": trim,
                        ("pairs" if kind == 1 else "primes"): products, "ms": round(ms, 4),
                        "fp64_equiv_tflops": round(flops / ms * 1e-9, 2),
                        "speedup_vs_native": round(entry["native_ms"] / ms, 3),
                        "speedup_vs_best_native": round(best_native / ms, 3),
                        "i8_kernel_ms": round(i8_ms, 4),
                        "i8_tops": round(i8["flops"] / max(i8["count"], 1) / i8_ms * 1e-9, 1),
                        "i8_frac_of_5pops": round(i8["flops"] / max(i8["count"], 1) / i8_ms * 1e-9 / 5000.0, 4),
                        "preprocess_ms": round(pre["ms"] / max(pre["count"], 1), 4),
                        "recombine_ms": round(rec["ms"] / max(rec["count"], 1), 4),
                        "errors": {}})
                else:
                    oz.gemm("N", "N", n, n, n, 1.0, da, n, db, n, 0.0, dc, n, stream)
                stream.sync()
                dc.download(out)
                oz.close()
                entry["configs"][cfg_i]["errors"]["e=%d" % e] = _sampled_error(A, B, out, rng)
            for x in (da, db, dc):
                x.free()
            del A, B, a, b
        results.append(entry)
    roc.close()
    return {
        "workload": "square DGEMM NN; A(i,k) = m 2^x_i, B(k,j) = m 2^y_j, m uniform +-[1,2), x, y uniform integers "
        "in [-e, e]; timing at e=10",
        "error": "normwise over 2048 sampled entries against 80-bit extended-precision dot products",
        "native": "own v_mfma_f64_16x16x4 kernel (ozaki_dgemm_native) and rocBLAS dgemm",
        "results": results,
    }


def _stack_sampled_reference(stack, a, bl, m, n, k, rng, count=2048):
    """flat C positions of `count` sampled elements of the C blocks the (sorted) stack addresses, and their
    values C = sum over the block's entries of A B as 80-bit extended-precision dot products (C zero)"""
    cofs = stack[:, 2] - 1
    blocks = np.unique(cofs)
    pick = blocks[rng.integers(0, len(blocks), count)]
    i, j = rng.integers(0, m, count), rng.integers(0, n, count)
    ref = np.zeros(count, dtype=np.longdouble)
    af = a.reshape(-1).astype(np.longdouble)
    bf = bl.reshape(-1).astype(np.longdouble)
    kk = np.arange(k)
    for c0 in blocks:
        ents = stack[cofs == c0]
        s = np.nonzero(pick == c0)[0]
        if len(s) == 0:
            continue
        arow = af[(ents[:, 0] - 1)[:, None, None] + i[s][None, :, None] + kk[None, None, :] * m]  # A(i, kk)
        bcol = bf[(ents[:, 1] - 1)[:, None, None] + kk[None, None, :] * n + j[s][None, :, None]]  # B(kk, j)
        ref[s] = (arow * bcol).sum(axis=(0, 2))
    return pick + i + j * m, ref


OZAKI_STACK_CONFIGS = tuple((s, s - 1) for s in range(3, 9)) + ((8, 0),)  # (slices, trim): triangles, all pairs


def ozaki_stack_section(lx, stream, stack_size=100000, reps=3):
    """Batched 64^3 stack (100k entries; pools of 2000 A, 2000 B, 1000 C blocks): native FP64
    (libsmm_acc_process) against ozaki_stack at slice counts 3 .. 8. Blocks hold mantissas uniform
    +-[1,2) times 2^x per A-block row and per B-block column, x uniform in [-e, e]."""
    from libxstream_amd.workload import make_stack_workload

    m = n = k = 64
    w = make_stack_workload(m, n, k, stack_size, nc=1000, na=2000, nb=2000, dtype=np.float64)
    rng = np.random.default_rng(20260701)
    cblk = (w.stack[:, 2] - 1) // (m * n)
    head = int(np.searchsorted(cblk, 40))  # the entries of the first 40 C blocks (sorted stack)
    flops = 2.0 * m * n * k * stack_size
    st = lx.DeviceBuffer.from_array(w.stack)
    zeros = np.zeros(w.c.size)
    entry = {"stack_size": stack_size, "m": m, "n": n, "k": k, "blocks_a_b_c": [2000, 2000, 1000], "configs": []}
    for e in (10, 0, 30):  # e = 10 first: it carries the timing
        sgn = lambda shape: np.where(rng.integers(0, 2, shape, dtype=np.int8) != 0, 1.0, -1.0)
        A = rng.uniform(1, 2, (2000, k, m)) * sgn((2000, k, m)) * np.exp2(rng.integers(-e, e + 1, (2000, 1, m)))
        Bh = rng.uniform(1, 2, (2000, n, k)) * sgn((2000, n, k)) * np.exp2(rng.integers(-e, e + 1, (2000, n, 1)))
        a = np.ascontiguousarray(A.reshape(2000, -1))
        bl = np.ascontiguousarray(Bh.transpose(0, 2, 1).reshape(2000, -1))  # kernel layout: k x n row-major
        da, db = lx.DeviceBuffer.from_array(a), lx.DeviceBuffer.from_array(bl)
        dc = lx.DeviceBuffer.from_array(zeros)
        out = np.empty(w.c.size)
        sel, ref = _stack_sampled_reference(w.stack[:head], a, bl, m, n, k, rng)
        err = lambda: float(np.sqrt(((out[sel].astype(np.longdouble) - ref) ** 2).sum() / (ref ** 2).sum()))
        if e == 10:
            ms = _event_ms(lx, stream, lambda: lx.process(st, stack_size, lx.REAL_8, da, db, dc, m, n, k, stream), reps)
            entry["native_ms"] = round(ms, 4)
            entry["native_tflops"] = round(flops / ms * 1e-9, 2)
            entry["native_errors"] = {}
        dc.upload(zeros)
        lx.process(st, stack_size, lx.REAL_8, da, db, dc, m, n, k, stream)
        stream.sync()
        dc.download(out)
        entry["native_errors"]["e=%d" % e] = err()
        for cfg_i, (sl, trim) in enumerate(OZAKI_STACK_CONFIGS):
            oz = lx.Ozaki(nslices=sl, trim=trim)
            if e == 10:
                ms = _event_ms(lx, stream, lambda: oz.stack(st, stack_size, m, n, k, da, 2000, db, 2000, dc, stream), reps)
                pairs = sum(1 for p in range(sl) for q in range(sl) if p + q <= 2 * (sl - 1) - trim)
                entry["configs"].append({"slices": sl, "trim": trim, "pairs": pairs, "ms": round(ms, 4),
                                         "fp64_equiv_tflops": round(flops / ms * 1e-9, 2),
                                         "speedup_vs_native": round(entry["native_ms"] / ms, 3), "errors": {}})
            dc.upload(zeros)
            oz.stack(st, stack_size, m, n, k, da, 2000, db, 2000, dc, stream)
            stream.sync()
            dc.download(out)
            oz.close()
            entry["configs"][cfg_i]["errors"]["e=%d" % e] = err()
        for x in (da, db, dc):
            x.free()
    st.free()
    entry["workload"] = ("64x64x64 stack of %d entries sorted by C (pools 2000/2000/1000 blocks); A blocks m 2^x per row, "
                         "B blocks m 2^y per column, m uniform +-[1,2), x, y uniform in [-e, e]; timing at e=10" % stack_size)
    entry["error"] = ("normwise over 2048 sampled elements of the first 40 C blocks against 80-bit extended-precision "
                      "dot products")
    return entry


def learned_section(lx):
    """The shipped predictor judged on the 100 held-out triplets per type by the exhaustive sweeps shipped
    with it (each method's choice looked up in the shape's measured configurations)."""
    import libxstream_amd.learned as learned

    params = os.path.join(os.path.dirname(lx.LIB_PATH), "params")
    model_path = os.path.join(params, "mlp_MI355X.txt")
    sweeps = [os.path.join(params, "sweep_%s_MI355X.jsonl.gz" % t) for t in ("f64", "f32")]
    if not os.path.exists(model_path) or not all(os.path.exists(x) for x in sweeps):
        return None
    sweep = learned.load_sweep(sweeps)
    train_t, held_t = learned.split(learned.triplets())
    tids = (lx.REAL_8, lx.REAL_4)
    train_keys = [(tid,) + t for tid in tids for t in train_t if (tid,) + t in sweep]
    held_keys = [(tid,) + t for tid in tids for t in held_t if (tid,) + t in sweep]
    model = lx.Mlp(path=model_path)
    t0 = time.perf_counter()
    results = learned.evaluate(model, sweep, held_keys, train_keys)
    t1 = time.perf_counter()
    model.close()
    rep = learned.report(results, tids)
    rep["workload"] = ("exhaustive sweeps (tune_multiply, S=30000) over 500 seeded triplets in 4..80; "
                       "400 train, 100 held out per type; fraction of the best measured configuration")
    rep["prediction_ms_per_shape"] = round((t1 - t0) * 1e3 / max(len(held_keys), 1), 3)
    return rep


def traffic_probe(args):
    """Child of rocprofv3 --pmc: one calibration copy, then three stack launches."""
    import libxstream_amd as lx

    w = make_workload(args)
    lx.init()
    stream = lx.Stream("probe")
    src = lx.DeviceBuffer(COPY_BYTES)
    dst = lx.DeviceBuffer(COPY_BYTES)
    src.zero(stream)
    stream.sync()
    lx.check(lx.lib().libxstream_mem_copy_d2d(src.ptr, dst.ptr, COPY_BYTES, stream.handle), "copy_d2d")
    stream.sync()
    res = Resident(lx, w, stream)
    for _ in range(3):
        res.process(stream)
    stream.sync()
    res.free()
    src.free()
    dst.free()
    stream.close()
    lx.finalize()


def pmc_run(args, counter, tmp):
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    out = os.path.join(tmp, counter.lower())
    cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", out, "-o", "pmc", "--",
           sys.executable, os.path.abspath(__file__), "--traffic-probe", "--m", str(args.m),
           "--stack-size", str(args.stack_size), "--dtype", args.dtype]
    if args.n:
        cmd += ["--n", str(args.n)]
    if args.k:
        cmd += ["--k", str(args.k)]
    subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=180, check=False)
    rows = []
    for path in glob.glob(os.path.join(out, "**", "*counter_collection*.csv"), recursive=True):
        with open(path) as f:
            rows += list(csv.DictReader(f))
    copy_vals, smm_vals = [], []
    for r in rows:
        if r.get("Counter_Name") != counter:
            continue
        name = r.get("Kernel_Name", "")
        val = float(r.get("Counter_Value", 0))
        if "smm_" in name:
            smm_vals.append(val)
        elif "opy" in name:  # the runtime's device-to-device copy kernel
            copy_vals.append(val)
    if not copy_vals or not smm_vals:
        return None
    # counter unit per byte, from the copy of known size (reads and writes COPY_BYTES each)
    factor = max(copy_vals) / COPY_BYTES
    return {"per_launch_bytes": float(np.median(smm_vals)) / factor, "copy_counter": max(copy_vals),
            "factor_per_byte": factor}


def traffic(args):
    try:
        with tempfile.TemporaryDirectory(prefix="lxs_pmc_") as tmp:
            fetch = pmc_run(args, "FETCH_SIZE", tmp)
            write = pmc_run(args, "WRITE_SIZE", tmp)
    except (OSError, subprocess.SubprocessError):
        return None, None
    if fetch is None or write is None:
        return None, None
    return fetch["per_launch_bytes"] + write["per_launch_bytes"], {
        "fetch_bytes": round(fetch["per_launch_bytes"]), "write_bytes": round(write["per_launch_bytes"]),
        "fetch_counter_per_byte": fetch["factor_per_byte"], "write_counter_per_byte": write["factor_per_byte"],
        "calibration": "device-to-device copy of %d bytes in the same run" % COPY_BYTES}


def cpu_baseline(w, seconds):
    """Reference host path timed on this host (oracle/cpu_baseline.c), bounded to ~`seconds`."""
    import oracle

    nthreads = max(1, min(16, os.cpu_count() or 1))
    sample = np.ascontiguousarray(w.stack)  # the same stack; repeated until ~`seconds` of CPU work
    c = np.zeros_like(w.c)
    t1 = oracle.cpu_stack(w.m, w.n, w.k, sample, w.a, w.b, c, nrepeat=1, warmup=1, nthreads=nthreads)
    nrep = int(max(1, min(1000, seconds / max(t1, 1e-6))))
    t = oracle.cpu_stack(w.m, w.n, w.k, sample, w.a, w.b, c, nrepeat=nrep, warmup=0, nthreads=nthreads)
    flops = 2.0 * w.m * w.n * w.k * len(sample) * nrep
    return {
        "value": round(flops / t * 1e-9, 2),
        "unit": "GFLOP/s",
        "cores": nthreads,
        "kind": "port",
        "sample": "the full %d-entry stack x %d passes (%.1f s), B column-major as the reference host path"
        % (len(sample), nrep, t),
    }


def main():
    args = parse()
    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (the stack path shards by C-run "
              "ranges with no exchange between GPUs)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.sweep:
        import libxstream_amd as lx

        lx.init()
        stream = lx.Stream("bench")
        rows = sweep_section(lx, stream)
        stream.close()
        lx.finalize()
        print(json.dumps({"sweep": rows}))
        return
    if args.ozaki_only:
        import libxstream_amd as lx

        lx.init()
        stream = lx.Stream("bench")
        out = ozaki_section(lx, stream, [int(x) for x in args.ozaki_sizes.split(",") if x])
        out["stack_64"] = ozaki_stack_section(lx, stream)
        stream.close()
        lx.finalize()
        print(json.dumps({"ozaki": out}))
        return
    # PMC child runs first, before this process initialises the GPU
    tr, tr_detail = (None, None) if args.no_traffic else traffic(args)

    import libxstream_amd as lx
    from libxstream_amd.workload import make_stack_workload

    w = make_workload(args)
    m, n, k = w.m, w.n, w.k
    lx.init()
    stream = lx.Stream("bench")
    res = Resident(lx, w, stream)
    cfg, src = lx.config_get(res.tid, m, n, k)
    ms_per_step, kernel_ms = timed(lx, res, stream, args.steps, args.warmup)
    trans = transpose_section(lx, res, stream)
    res.free()
    copyin = copyin_section(lx, w, stream)
    small = None
    if not args.no_small:  # the 30000-entry stack over 2000/2000/1000 blocks (the reference's stack size)
        ws = make_stack_workload(m, n, k, 30000, nc=1000, na=2000, nb=2000, dtype=w.dtype, seed=25071975)
        rs = Resident(lx, ws, stream)
        s_step, s_kernel = timed(lx, rs, stream, max(args.steps, 20), args.warmup)
        rs.free()
        small = {"stack_size": 30000, "blocks_a_b_c": [2000, 2000, 1000], "ms_per_step": round(s_step, 5),
                 "gflops": round(ws.flops() / (s_step * 1e-3) * 1e-9, 1), "kernel_ms": round(s_kernel, 5),
                 "kernel_gflops": round(ws.flops() / (s_kernel * 1e-3) * 1e-9, 1)}
    transfer = transfer_section(lx, m, n, k, w.dtype, stream)
    batch = None if args.no_batch else batched(lx, stream, args.steps, args.warmup)
    ozaki = None if args.no_ozaki else ozaki_section(lx, stream, [int(x) for x in args.ozaki_sizes.split(",") if x])
    if ozaki is not None:
        ozaki["stack_64"] = ozaki_stack_section(lx, stream)
    learned_rep = None if args.no_learned else learned_section(lx)
    ncus, arch = lx.device_info()
    stream.close()
    lx.finalize()

    flops = w.flops()
    value = flops / (ms_per_step * 1e-3) * 1e-9
    bytes_alg = w.bytes_alg()
    achieved = bytes_alg / (kernel_ms * 1e-3) * 1e-9
    cpu = None
    if not args.no_cpu and args.cpu_seconds > 0:
        cpu = cpu_baseline(w, args.cpu_seconds)

    line = {
        "metric": "SMM stack GFLOP/s (FP64 23^3)" if (m, n, k, args.dtype) == (23, 23, 23, "f64")
        else "SMM stack GFLOP/s (%s %dx%dx%d)" % (args.dtype.upper(), m, n, k),
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic (seeded uniform[-1,1] pools, sorted-by-C stack, runs 30+-4)",
        "config": {
            "workload": "libsmm_acc_process %dx%dx%d %s stack S=%d (C blocks %d, A/B blocks %d/%d), "
            "1-based offsets sorted by C" % (m, n, k, args.dtype, w.stack_size, w.nc, w.na, w.nb),
            "m": m, "n": n, "k": k, "stack_size": w.stack_size,
            "kernel_config": {kk: v for kk, v in cfg.as_dict().items() if kk != "gflops"},
            "config_source": ["default", "tuned", "env", "learned"][src],
            "device": "%s (%d CUs)" % (arch, ncus),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": (round(tr) if tr is not None else None),
            "traffic_detail": tr_detail,
            # the physical bytes behind L2 per launch (PMC) over the same kernel time, against the same peak
            "traffic_gbs": (round(tr / (kernel_ms * 1e-3) * 1e-9, 1) if tr is not None else None),
            "traffic_frac": (round(tr / (kernel_ms * 1e-3) * 1e-9 / HBM_PEAK_GBS, 4) if tr is not None else None),
            "kernel": "smm_mfma (libsmm_acc_process)",
            "kernel_ms": round(kernel_ms, 5),
            "bytes_alg_per_launch": bytes_alg,
            "flops_per_launch": flops,
            "fp64_tflops": round(flops / (kernel_ms * 1e-3) * 1e-12, 3),
            "fp64_frac": round(flops / (kernel_ms * 1e-3) * 1e-12 / FP64_PEAK_TFLOPS, 4),
        },
        "cpu_baseline": cpu,
        "stack_30000": small,
        "transpose_b_pool": trans,
        "copy_in_pinned": copyin,
        "stack_transfer": transfer,
        "batched_mixed": batch,
        "ozaki": ozaki,
        "learned": learned_rep,
        "reference_published_gflops_23cubed_fp64_s30000": {"V100": 2459.7, "P100": 1657.8},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
