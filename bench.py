#!/usr/bin/env python3
"""Benchmark: DBCSR stack multiply (libsmm_acc_process) on one MI355X, FP64 23x23x23.

Workload (the configuration the metric is quoted on): m = n = k = 23, FP64, pools of 2000 A and 2000 B
blocks (uniform[-1,1], seeded) and 1000 C blocks (zero), a parameter stack of 30000 (a, b, c) 1-based
element offsets sorted by C (runs of 30 +- 4), a and b uniform. One step = one libsmm_acc_process call
over the whole stack with inputs resident in HBM, like one repetition of samples/smm/acc_bench.c:338-342.

The LAST stdout line is one compact JSON record (about 2 KB, the shape of acc_bench.c:349-350's
closing report line):
  value         whole-job GFLOP/s = 2*m*n*k*S / host wall time per step (stream-synchronised)
  roofline      the stack kernel's algorithmic bytes S*8*(mk+kn+2mn) (samples/smm/smm_kernel.c:504-506)
                per launch / its average launch duration (one pair of HIP events on the launch stream
                around the timed launches), against 8 TB/s; `traffic` = bytes per launch behind L2 from
                rocprofv3 PMC counters (FETCH_SIZE, WRITE_SIZE) in child runs, each scaled by what the
                same counter reads for a device-to-device copy of known size; `mfma_fp64_tflops` from
                SQ_INSTS_VALU_MFMA_MOPS_F64 (512 flops per unit) over the same launch duration
  cpu_baseline  the reference host path (OpenMP, whole C-runs per thread, AVX-512/AVX2 entry kernel) on
                all of the host's cores, timed on the same stack
  secondary     the 500000-entry stack, the batched mixed-shape graph, Ozaki-int8 DGEMM against rocBLAS,
                and the learned predictor, each summarised in one short object
Everything else (sweeps, per-stack details, transfer, transpose) goes to the --report JSON file.
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
METRIC = "SMM stack GFLOP/s (FP64 23\u00b3) & Ozaki-int8/native-FP64 speedup at equal error"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--m", type=int, default=23)
    p.add_argument("--n", type=int, default=None)
    p.add_argument("--k", type=int, default=None)
    p.add_argument("--stack-size", type=int, default=30000)
    p.add_argument("--nc", type=int, default=None, help="C blocks (default 1000 at S=30000, else S/30)")
    p.add_argument("--na", type=int, default=None, help="A blocks (default 2000 at S=30000, else 2 nc)")
    p.add_argument("--nb", type=int, default=None, help="B blocks (default 2000 at S=30000, else 2 nc)")
    p.add_argument("--dtype", choices=["f64", "f32"], default="f64")
    p.add_argument("--cpu-seconds", type=float, default=10.0, help="target CPU-baseline time (0 = skip)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-traffic", action="store_true", help="skip the rocprofv3 PMC child runs")
    p.add_argument("--no-extras", action="store_true", help="only the headline stack (no secondary sections)")
    p.add_argument("--no-large", action="store_true", help="skip the 500000-entry stack")
    p.add_argument("--no-batch", action="store_true", help="skip the batched mixed-shape measurement")
    p.add_argument("--no-ozaki", action="store_true", help="skip the Ozaki-int8 DGEMM measurement")
    p.add_argument("--ozaki-sizes", default="1024,4096,8192", help="square DGEMM sizes for the Ozaki section")
    p.add_argument("--ozaki-only", action="store_true", help="print only the Ozaki section (no stack measurement)")
    p.add_argument("--no-learned", action="store_true", help="skip the learned-predictor evaluation")
    p.add_argument("--report", default=os.path.join("gpurun_out", "bench_report.json"),
                   help="JSON file for the full secondary sections ('' = none)")
    p.add_argument("--sweep", action="store_true",
                   help="print only the plain-stack sweep: 7 shapes x FP64/FP32 x S in {30k, 100k, 500k}")
    p.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    return p.parse_args()


def pools(args):
    """(nc, na, nb): 1000/2000/2000 at S = 30000, else S/30 C blocks and twice as many A and B blocks"""
    if args.stack_size == 30000:
        nc, na, nb = 1000, 2000, 2000
    else:
        nc = max(1, int(round(args.stack_size / 30.0)))
        na = nb = 2 * nc
    return (args.nc or nc, args.na or na, args.nb or nb)


def make_workload(args):
    from libxstream_amd.workload import make_stack_workload

    m = args.m
    n = args.n or m
    k = args.k or m
    dtype = np.float64 if args.dtype == "f64" else np.float32
    nc, na, nb = pools(args)
    return make_stack_workload(m, n, k, args.stack_size, nc=nc, na=na, nb=nb, dtype=dtype, seed=25071975)


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


def timed(lx, res, stream, steps, warmup, settle_s=0.2):
    """(host ms per step, ms per launch between two HIP events on the launch stream around the same
    timed launches). Before the warmup steps the stack runs untimed for settle_s seconds, so that the
    device clocks have left their idle state when the timed steps begin."""
    t_end = time.perf_counter() + settle_s
    while time.perf_counter() < t_end:
        for _ in range(10):
            res.process(stream)
        stream.sync()
    for _ in range(warmup):
        res.process(stream)
    stream.sync()
    e0, e1 = lx.Event(), lx.Event()
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        res.process(stream)
    e1.record(stream)
    stream.sync()
    t1 = time.perf_counter()
    kernel_ms = e0.elapsed_ms(e1) / steps
    e0.close()
    e1.close()
    return (t1 - t0) * 1e3 / steps, kernel_ms


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
    lx.profile_enable(False)
    # every schedule kind replayed the same way: the persistent kernel (default and tuned chunking), the
    # per-stack launches in order, and the per-stack launches on concurrent lanes
    schedules = {}
    for name, params in (("persistent_default", lx.BatchParams(0, 0, 0)), ("per_stack_in_order", lx.BatchParams(0, 0, 1)),
                         ("per_stack_lanes", lx.BatchParams(0, 0, 2))):
        bb = lx.Batch(lx.REAL_8, items, params)
        for _ in range(warmup):
            bb.replay(stream)
        stream.sync()
        ta = time.perf_counter()
        for _ in range(steps):
            bb.replay(stream)
        stream.sync()
        schedules[name] = round((time.perf_counter() - ta) * 1e3 / steps, 4)
        bb.close()
    lx.profile_enable(True)
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
        "graph_ms_by_schedule": schedules,
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


def _two_prod(x, y):
    """x*y = p + e exactly, elementwise (Dekker's product with Veltkamp's splitting; numpy does not contract)"""
    p = x * y
    t = 134217729.0 * x  # 2^27 + 1
    xh = t - (t - x)
    xl = x - xh
    t = 134217729.0 * y
    yh = t - (t - y)
    yl = y - yh
    return p, ((xh * yh - p) + xh * yl + xl * yh) + xl * yl


def _dd_rows(p, e):
    """double-double sums (hi, lo) of the rows of p + e: hi the correctly rounded sum of the exact terms
    (math.fsum), lo the correctly rounded remainder"""
    import math

    hi, lo = np.empty(p.shape[0]), np.empty(p.shape[0])
    for r in range(p.shape[0]):
        terms = p[r].tolist() + e[r].tolist()
        hi[r] = math.fsum(terms)
        terms.append(-hi[r])
        lo[r] = math.fsum(terms)
    return hi, lo


def _gemm_reference(A, B, rng, count=1024):
    """(flat column-major positions, hi, lo) of `count` sampled entries of A B in double-double"""
    m, n = A.shape[0], B.shape[1]
    i, j = rng.integers(0, m, count), rng.integers(0, n, count)
    hi, lo = np.empty(count), np.empty(count)
    for s0 in range(0, count, 128):
        p, e = _two_prod(A[i[s0:s0 + 128]], np.ascontiguousarray(B[:, j[s0:s0 + 128]].T))
        hi[s0:s0 + 128], lo[s0:s0 + 128] = _dd_rows(p, e)
    return i + j * m, hi, lo


def _ref_error(ref, c_flat):
    """normwise error ||C - C_dd|| / ||C_dd|| over the sampled entries of a double-double reference"""
    idx, hi, lo = ref
    d = (c_flat[idx] - hi) - lo
    return float(np.sqrt((d * d).sum() / (hi * hi).sum()))


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
            ref = _gemm_reference(A, B, rng)
            if e == 10:  # timing (independent of the exponent range)
                ms = _event_ms(lx, stream, lambda: lx.dgemm_native("N", "N", n, n, n, 1.0, da, n, db, n, 0.0, dc, n,
                                                                   stream), reps)
                dc.download(out)
                entry["native_ms"] = round(ms, 4)
                entry["native_tflops"] = round(flops / ms * 1e-9, 2)
                entry["native_error"] = _ref_error(ref, out)
                if roc.lib is not None:
                    ms = _event_ms(lx, stream, lambda: roc.dgemm(n, da, db, dc), reps)
                    dc.download(out)
                    entry["rocblas_ms"] = round(ms, 4)
                    entry["rocblas_tflops"] = round(flops / ms * 1e-9, 2)
                    entry["rocblas_error"] = _ref_error(ref, out)
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
                entry["configs"][cfg_i]["errors"]["e=%d" % e] = _ref_error(ref, out)
            for x in (da, db, dc):
                x.free()
            del A, B, a, b
        results.append(entry)
    roc.close()
    return {
        "workload": "square DGEMM NN; A(i,k) = m 2^x_i, B(k,j) = m 2^y_j, m uniform +-[1,2), x, y uniform integers "
        "in [-e, e]; timing at e=10",
        "error": "normwise over 1024 sampled entries against double-double dot products (exact products, "
        "correctly rounded sums)",
        "native": "own v_mfma_f64_16x16x4 kernel (ozaki_dgemm_native) and rocBLAS dgemm",
        "results": results,
    }


def _stack_sampled_reference(stack, a, bl, m, n, k, rng, count=1024):
    """(flat C positions, hi, lo) of `count` sampled elements of the C blocks the (sorted) stack addresses:
    C = sum over the block's entries of A B in double-double (C zero)"""
    cofs = stack[:, 2] - 1
    blocks = np.unique(cofs)
    pick = blocks[rng.integers(0, len(blocks), count)]
    i, j = rng.integers(0, m, count), rng.integers(0, n, count)
    hi, lo = np.zeros(count), np.zeros(count)
    af = a.reshape(-1)
    bf = bl.reshape(-1)
    kk = np.arange(k)
    for c0 in blocks:
        ents = stack[cofs == c0]
        s = np.nonzero(pick == c0)[0]
        if len(s) == 0:
            continue
        arow = af[(ents[:, 0] - 1)[:, None, None] + i[s][None, :, None] + kk[None, None, :] * m]  # A(i, kk)
        bcol = bf[(ents[:, 1] - 1)[:, None, None] + kk[None, None, :] * n + j[s][None, :, None]]  # B(kk, j)
        p, e = _two_prod(arow.transpose(1, 0, 2).reshape(len(s), -1), bcol.transpose(1, 0, 2).reshape(len(s), -1))
        hi[s], lo[s] = _dd_rows(p, e)
    return pick + i + j * m, hi, lo


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
        ref = _stack_sampled_reference(w.stack[:head], a, bl, m, n, k, rng)
        err = lambda: _ref_error(ref, out)
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
    entry["error"] = ("normwise over 1024 sampled elements of the first 40 C blocks against double-double sums "
                      "(exact products, correctly rounded sums)")
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


def pmc_run(args, counters, tmp):
    """{counter: (median value per stack launch, largest value of the calibration copy)} from one
    rocprofv3 --pmc child run of traffic_probe, or None"""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    out = os.path.join(tmp, "_".join(c.lower() for c in counters))
    cmd = [prof, "--pmc"] + list(counters) + ["--output-format", "csv", "-d", out, "-o", "pmc", "--",
                                              sys.executable, os.path.abspath(__file__), "--traffic-probe",
                                              "--m", str(args.m), "--stack-size", str(args.stack_size),
                                              "--dtype", args.dtype]
    for flag in ("n", "k", "nc", "na", "nb"):
        if getattr(args, flag):
            cmd += ["--" + flag, str(getattr(args, flag))]
    subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=120, check=False)
    rows = []
    for path in glob.glob(os.path.join(out, "**", "*counter_collection*.csv"), recursive=True):
        with open(path) as f:
            rows += list(csv.DictReader(f))
    result = {}
    for counter in counters:
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
        if smm_vals:
            result[counter] = (float(np.median(smm_vals)), max(copy_vals) if copy_vals else None)
    return result or None


def traffic(args):
    """PMC figures of the stack kernel per launch: HBM-side bytes (FETCH_SIZE, WRITE_SIZE, each calibrated
    by a device copy of COPY_BYTES that reads and writes that many bytes) and FP64 matrix-core work
    (SQ_INSTS_VALU_MFMA_MOPS_F64 in units of 512 flops, SQ_VALU_MFMA_BUSY_CYCLES, GRBM_GUI_ACTIVE)"""
    out = {}
    try:
        with tempfile.TemporaryDirectory(prefix="lxs_pmc_") as tmp:
            for counters in (("FETCH_SIZE",), ("WRITE_SIZE",),
                             ("SQ_INSTS_VALU_MFMA_MOPS_F64", "SQ_VALU_MFMA_BUSY_CYCLES", "GRBM_GUI_ACTIVE")):
                got = pmc_run(args, counters, tmp)
                if got:
                    out.update(got)
    except (OSError, subprocess.SubprocessError):
        pass
    return out


def cpu_baseline(w, seconds):
    """Reference host path timed on this host (oracle/cpu_baseline.c) on the same stack, on every core
    this process may run on (and on 16 threads when that is fewer, keeping the faster), bounded to
    about `seconds` of CPU work in total."""
    import oracle

    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    best = None
    for nthreads in sorted({avail, min(16, avail)}, reverse=True):
        c = np.zeros_like(w.c)
        t1 = oracle.cpu_stack(w.m, w.n, w.k, w.stack, w.a, w.b, c, nrepeat=1, warmup=1, nthreads=nthreads)
        nrep = int(max(1, min(10000, 0.5 * seconds / max(t1, 1e-6))))
        t = oracle.cpu_stack(w.m, w.n, w.k, w.stack, w.a, w.b, c, nrepeat=nrep, warmup=0, nthreads=nthreads)
        gflops = 2.0 * w.m * w.n * w.k * w.stack_size * nrep / t * 1e-9
        if best is None or gflops > best[0]:
            best = (gflops, nthreads, nrep, t)
    gflops, nthreads, nrep, t = best
    return {
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "cores": nthreads,
        "kind": "port",
        "sample": "the same %d-entry stack x %d passes (%.1f s), B column-major as the reference host path; "
        "%s entry kernel; %d cores available" % (w.stack_size, nrep, t, oracle.cpu_isa(), avail),
    }


def _compact_ozaki(oz):
    """best Ozaki-int8 configuration at <= 1e-14 per DGEMM size, against rocBLAS and our FP64 kernel"""
    out = []
    for entry in oz["results"]:
        ok = [c for c in entry["configs"] if all(v is not None and v <= 1e-14 for v in c["errors"].values())]
        if not ok:
            continue
        best = min(ok, key=lambda c: c["ms"])
        # scheme 2 reports the moduli it used (at most max_primes), scheme 1 its slices
        used = best.get("primes") or best.get("max_primes") or best.get("slices")
        out.append({"n": entry["n"], "scheme": best["scheme"], "moduli_or_slices": used, "trim": best["trim"],
                    "ms": best["ms"], "rocblas_ms": entry.get("rocblas_ms"),
                    "vs_rocblas": round(entry["rocblas_ms"] / best["ms"], 3) if entry.get("rocblas_ms") else None,
                    "vs_own_fp64": best["speedup_vs_native"], "max_error": float("%.2g" % max(best["errors"].values()))})
    return out


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
    pmc = {} if args.no_traffic else traffic(args)
    print("bench: PMC counters %s" % sorted(pmc), file=sys.stderr, flush=True)

    import libxstream_amd as lx
    from libxstream_amd.workload import make_stack_workload

    w = make_workload(args)
    m, n, k = w.m, w.n, w.k
    lx.init()
    stream = lx.Stream("bench")
    res = Resident(lx, w, stream)
    cfg, src = lx.config_get(res.tid, m, n, k)
    ms_per_step, kernel_ms = timed(lx, res, stream, args.steps, args.warmup)
    print("bench: %dx%dx%d S=%d %.4f ms/step, %.4f ms/launch" % (m, n, k, w.stack_size, ms_per_step, kernel_ms),
          file=sys.stderr, flush=True)
    report = {}
    secondary = {}
    extras = not args.no_extras
    if extras:
        report["transpose_b_pool"] = transpose_section(lx, res, stream)
    res.free()
    if extras:
        report["copy_in_pinned"] = copyin_section(lx, w, stream)
        report["stack_transfer"] = transfer_section(lx, m, n, k, w.dtype, stream)
    if extras and not args.no_large:  # the 500000-entry stack (pools of 16667/33334/33334 blocks: beyond the MALL)
        wl = make_stack_workload(m, n, k, 500000, dtype=w.dtype, seed=25071975)
        rl = Resident(lx, wl, stream)
        l_step, l_kernel = timed(lx, rl, stream, max(5, args.steps // 5), 2)
        rl.free()
        secondary["stack_500000"] = {
            "pools_a_b_c": [wl.na, wl.nb, wl.nc], "ms_per_step": round(l_step, 4), "kernel_ms": round(l_kernel, 4),
            "gflops": round(wl.flops() / (l_step * 1e-3) * 1e-9, 1),
            "alg_frac": round(wl.bytes_alg() / (l_kernel * 1e-3) * 1e-9 / HBM_PEAK_GBS, 3)}
        print("bench: S=500000 %.4f ms/launch" % l_kernel, file=sys.stderr, flush=True)
        del wl, rl
    if extras and not args.no_batch:
        report["batched_mixed"] = b = batched(lx, stream, 20, 3)
        secondary["batched_mixed"] = {"stacks": 24, "entries": b["entries"], "graph_ms": b["graph_ms_per_step"],
                                      "graph_gflops": b["graph_gflops"], "schedule": {0: "persistent kernel", 1: "per-stack launches",
                                                   2: "per-stack launches on 4 concurrent lanes"}[
                                          b["schedule"].get("per_stack") or 0],
                                      "separate_launches_ms": b["separate_launches_ms_per_step"],
                                      "graph_ms_by_schedule": b["graph_ms_by_schedule"]}
        print("bench: batched %.4f ms" % b["graph_ms_per_step"], file=sys.stderr, flush=True)
    if extras and not args.no_ozaki:
        report["ozaki"] = oz = ozaki_section(lx, stream, [int(x) for x in args.ozaki_sizes.split(",") if x])
        oz["stack_64"] = ozs = ozaki_stack_section(lx, stream)
        ok = [c for c in ozs["configs"] if all(v <= 1e-14 for v in c["errors"].values())]
        secondary["ozaki_dgemm"] = _compact_ozaki(oz)
        if ok:
            best = min(ok, key=lambda c: c["ms"])
            secondary["ozaki_stack_64"] = {"slices": best["slices"], "ms": best["ms"], "native_ms": ozs["native_ms"],
                                           "speedup": best["speedup_vs_native"],
                                           "max_error": float("%.2g" % max(best["errors"].values()))}
        print("bench: ozaki done", file=sys.stderr, flush=True)
    if extras and not args.no_learned:
        report["learned"] = rep = learned_section(lx)
        if rep is not None:
            secondary["learned_heldout"] = {f: {"mean": rep[f]["mean"], "within_10pct": rep[f]["within_10pct"]}
                                            for f in ("learned", "learned_top4", "default", "nearest")}
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
        print("bench: cpu baseline %s" % cpu, file=sys.stderr, flush=True)

    traffic_bytes = mfma = None
    if "FETCH_SIZE" in pmc and "WRITE_SIZE" in pmc and pmc["FETCH_SIZE"][1] and pmc["WRITE_SIZE"][1]:
        fetch = pmc["FETCH_SIZE"][0] / (pmc["FETCH_SIZE"][1] / COPY_BYTES)
        write = pmc["WRITE_SIZE"][0] / (pmc["WRITE_SIZE"][1] / COPY_BYTES)
        traffic_bytes = fetch + write
    if "SQ_INSTS_VALU_MFMA_MOPS_F64" in pmc:
        mfma = 512.0 * pmc["SQ_INSTS_VALU_MFMA_MOPS_F64"][0]
    roof = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": round(traffic_bytes) if traffic_bytes is not None else None,
        "traffic_frac": round(traffic_bytes / (kernel_ms * 1e-3) * 1e-9 / HBM_PEAK_GBS, 4) if traffic_bytes else None,
        "kernel": "smm_mfma (libsmm_acc_process)", "kernel_ms": round(kernel_ms, 5),
        "bytes_alg_per_launch": bytes_alg,
        "fp64_tflops": round(flops / (kernel_ms * 1e-3) * 1e-12, 3),
        "fp64_frac": round(flops / (kernel_ms * 1e-3) * 1e-12 / FP64_PEAK_TFLOPS, 4),
        "mfma_fp64_tflops": round(mfma / (kernel_ms * 1e-3) * 1e-12, 3) if mfma else None,
        # SQ_VALU_MFMA_BUSY_CYCLES sums over the SIMDs (4 per CU), GRBM_GUI_ACTIVE over the XCDs (32 CUs each)
        "mfma_busy": (round(pmc["SQ_VALU_MFMA_BUSY_CYCLES"][0] / (4.0 * ncus * pmc["GRBM_GUI_ACTIVE"][0] / max(1, ncus // 32)), 4)
                      if "SQ_VALU_MFMA_BUSY_CYCLES" in pmc and pmc.get("GRBM_GUI_ACTIVE", (0,))[0] else None),
    }
    report["pmc"] = {c: {"per_launch": v[0], "copy_of_%d_bytes" % COPY_BYTES: v[1]} for c, v in pmc.items()}
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic: seeded uniform[-1,1] pools, zero C, sorted-by-C stack with runs of 30+-4",
        "config": {
            "workload": "libsmm_acc_process %dx%dx%d %s, S=%d, pools A/B/C %d/%d/%d blocks, 1-based offsets "
            "sorted by C" % (m, n, k, args.dtype, w.stack_size, w.na, w.nb, w.nc),
            "kernel_config": {kk: v for kk, v in cfg.as_dict().items() if kk not in ("gflops",)},
            "config_source": ["default", "tuned", "env", "learned"][src],
            "device": "%s (%d CUs)" % (arch, ncus),
        },
        "roofline": roof,
        "cpu_baseline": cpu,
        "secondary": secondary,
        "report": args.report or None,
    }
    if args.report:
        os.makedirs(os.path.dirname(os.path.abspath(args.report)), exist_ok=True)
        with open(args.report, "w") as f:
            json.dump(dict(report, headline=line), f)
    print(json.dumps(line, separators=(",", ":")))


if __name__ == "__main__":
    main()
