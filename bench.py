#!/usr/bin/env python3
"""Benchmark: DBCSR stack multiply (libsmm_acc_process) on one MI355X, FP64 23x23x23.

One step = one libsmm_acc_process call over the whole parameter stack (inputs resident in HBM),
like one repetition of samples/smm/acc_bench.c:338-342. Prints one JSON line.

  value         whole-job GFLOP/s = 2*m*n*k*S / wall time per step (host clock, stream-synchronised)
  roofline      the stack kernel's algorithmic bytes S*8*(mk+kn+2mn) (samples/smm/smm_kernel.c:504-506)
                per launch / its average duration from HIP events recorded on the launch stream,
                against 8 TB/s HBM
  cpu_baseline  the reference host path (OpenMP, whole C-runs per thread) on the host's cores,
                timed on a bounded sample of the same stack
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


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
    return p.parse_args()


def cpu_baseline(w, seconds):
    """Reference host path timed on this host (oracle/cpu_baseline.c), bounded to ~`seconds`."""
    import oracle

    nthreads = max(1, min(16, os.cpu_count() or 1))
    # bounded sample: the first min(S, 100000) entries of the same stack, repeated to ~`seconds`
    sample = np.ascontiguousarray(w.stack[: min(w.stack_size, 100000)])
    c = np.zeros_like(w.c)
    t1 = oracle.cpu_stack(w.m, w.n, w.k, sample, w.a, w.b, c, nrepeat=1, warmup=1, nthreads=nthreads)
    nrep = int(max(1, min(200, seconds / max(t1, 1e-6))))
    t = oracle.cpu_stack(w.m, w.n, w.k, sample, w.a, w.b, c, nrepeat=nrep, warmup=0, nthreads=nthreads)
    flops = 2.0 * w.m * w.n * w.k * len(sample) * nrep
    return {
        "value": round(flops / t * 1e-9, 2),
        "unit": "GFLOP/s",
        "cores": nthreads,
        "kind": "port",
        "sample": "%d-entry prefix of the same stack x %d passes (%.1f s), B column-major as the reference host path"
        % (len(sample), nrep, t),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported (the stack path shards by C-run "
              "ranges with no exchange, see DESIGN.md)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import libxstream_amd as lx
    from libxstream_amd.workload import make_stack_workload

    m = args.m
    n = args.n or m
    k = args.k or m
    dtype = np.float64 if args.dtype == "f64" else np.float32
    tid = lx.typeid(dtype)
    w = make_stack_workload(m, n, k, args.stack_size, dtype=dtype, seed=25071975)

    lx.init()
    stream = lx.Stream("bench")
    a = lx.DeviceBuffer.from_array(w.a)
    b = lx.DeviceBuffer.from_array(w.b)
    trs = lx.DeviceBuffer.from_array(np.arange(w.nb, dtype=np.int32) * (k * n))
    c = lx.DeviceBuffer.from_array(w.c)
    st = lx.DeviceBuffer.from_array(w.stack)
    lx.transpose(trs, 0, w.nb, b, tid, k, n, stream)  # B to the kernel's row-major layout (acc_bench.c:325)
    cfg, src = lx.config_get(tid, m, n, k)
    for _ in range(args.warmup):
        assert lx.process(st, w.stack_size, tid, a, b, c, m, n, k, stream) == 0
    stream.sync()

    lx.profile_enable(True)
    lx.profile_reset()
    stream.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lx.process(st, w.stack_size, tid, a, b, c, m, n, k, stream)
    stream.sync()
    t1 = time.perf_counter()
    prof = lx.profile_get("smm")
    lx.profile_enable(False)

    ms_per_step = (t1 - t0) * 1e3 / args.steps
    flops = w.flops()
    value = flops / (ms_per_step * 1e-3) * 1e-9
    kernel_ms = prof["ms"] / max(prof["count"], 1)
    bytes_alg = w.bytes_alg()
    achieved = bytes_alg / (kernel_ms * 1e-3) * 1e-9
    ncus, arch = lx.device_info()

    cpu = None
    if not args.no_cpu and args.cpu_seconds > 0:
        cpu = cpu_baseline(w, args.cpu_seconds)

    for buf in (a, b, trs, c, st):
        buf.free()
    stream.close()
    lx.finalize()

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
            "config_source": src,
            "device": "%s (%d CUs)" % (arch, ncus),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "smm_stack",
            "kernel_ms": round(kernel_ms, 5),
            "bytes_alg_per_launch": bytes_alg,
            "flops_per_launch": flops,
            "fp64_tflops": round(flops / (kernel_ms * 1e-3) * 1e-12, 3),
            "fp64_frac": round(flops / (kernel_ms * 1e-3) * 1e-12 / FP64_PEAK_TFLOPS, 4),
        },
        "cpu_baseline": cpu,
        "reference_published_gflops_23cubed_fp64": {"V100": 2459.7, "P100": 1657.8},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
