"""Shared helpers for the GPU tests: run a stack workload through the C ABI, compare with the oracle."""
import os

import numpy as np

import libxstream_amd as lx
import oracle

TOL = {np.dtype(np.float64): 1e-13, np.dtype(np.float32): 1e-5}


def tol(dtype):
    return TOL[np.dtype(dtype)]


def run_stack(stream, w, config=None, device_transpose=True, stack=None, repeat=1):
    """Upload w, (transpose B on the device), apply the stack `repeat` times, return C."""
    tid = lx.typeid(w.dtype)
    stack = w.stack if stack is None else stack
    bufs = []
    try:
        a = lx.DeviceBuffer.from_array(w.a)
        bufs.append(a)
        if device_transpose:
            b = lx.DeviceBuffer.from_array(w.b)
            trs = lx.DeviceBuffer.from_array(np.arange(w.nb, dtype=np.int32) * (w.k * w.n))
            bufs += [b, trs]
            lx.transpose(trs, 0, w.nb, b, tid, w.k, w.n, stream)
        else:
            b = lx.DeviceBuffer.from_array(w.b_device_layout())
            bufs.append(b)
        c = lx.DeviceBuffer.from_array(w.c)
        st = lx.DeviceBuffer.from_array(np.ascontiguousarray(stack, dtype=np.int32))
        bufs += [c, st]
        for _ in range(repeat):
            if config is None:
                rc = lx.process(st, len(stack), tid, a, b, c, w.m, w.n, w.k, stream)
                assert rc == 0, "libsmm_acc_process returned %d" % rc
            else:
                lx.process_config(config, st, len(stack), tid, a, b, c, w.m, w.n, w.k, stream)
        stream.sync()
        out = np.empty_like(w.c)
        c.download(out)
        return out
    finally:
        for buf in bufs:
            buf.free()


def oracle_error(w, out, stack=None, repeat=1):
    stack = w.stack if stack is None else stack
    c_in = w.c.ravel()
    for _ in range(repeat - 1):
        hi, lo = oracle.stack_dd(w.m, w.n, w.k, stack, w.a, w.b, c_in=c_in)
        c_in = (hi + lo).astype(w.dtype)
    hi, lo = oracle.stack_dd(w.m, w.n, w.k, stack, w.a, w.b, c_in=c_in)
    return oracle.error(out.ravel(), hi, lo)


def checksum(w, stack=None):
    """sum over all C elements of the stack's update: sum_e (1^T A_e)(B_e 1), in float64."""
    stack = w.stack if stack is None else stack
    m, n, k = w.m, w.n, w.k
    a64 = w.a.astype(np.float64)
    b64 = w.b.astype(np.float64)
    colsum_a = a64.reshape(w.na, k, m).sum(axis=2)  # (na, k): column sums of each A block
    rowsum_b = b64.reshape(w.nb, n, k).sum(axis=1)  # (nb, k): row sums of each column-major B block
    ai = (stack[:, 0].astype(np.int64) - 1) // (m * k)
    bi = (stack[:, 1].astype(np.int64) - 1) // (k * n)
    total = np.einsum("ij,ij->i", colsum_a[ai], rowsum_b[bi])
    scale = np.einsum("ij,ij->i", np.abs(a64).reshape(w.na, k, m).sum(axis=2)[ai],
                      np.abs(b64).reshape(w.nb, n, k).sum(axis=1)[bi])
    return float(np.sum(total)), float(np.sum(scale))


def uncut_blocks(w, stack=None, limit=32):
    """C blocks whose run has at most `limit` entries (2 bs of the launch: never cut into pieces), so
    their bits do not depend on the order in which atomics land."""
    stack = w.stack if stack is None else stack
    runs = np.bincount((np.asarray(stack)[:, 2] - 1) // (w.m * w.n), minlength=w.nc)
    return runs <= limit


def cut_limit(tid, m, n, k, stack_size):
    """Longest C-run that libsmm_acc_process keeps whole for this shape and stack size: split x bs of the
    launch, bs scaled down for stacks shorter than the configuration's S (smm_kernel.c:463-473)."""
    cfg, _ = lx.config_get(tid, m, n, k)
    bs, split = (cfg.bs if cfg.bs > 0 else 32), cfg.split
    if cfg.bs_short > 0 and cfg.s > 0 and stack_size <= cfg.s:  # the configuration's short-stack pair
        bs, split = cfg.bs_short, cfg.split_short
    if bs > 1 and cfg.s > 1 and stack_size < cfg.s:
        bs = min(bs, max(1, -(-stack_size * max(bs, 32) // (cfg.s - 1))))
    # runs this short are never cut: 64 entries in stacks of at least 16384 entries unless the environment
    # sets the floor for every stack
    floor = int(os.environ.get("LIBXSTREAM_SMM_SPLIT_MIN") or 0) or (64 if stack_size >= 16384 else 1)
    return max((split if split > 0 else 2) * bs, floor)


def sampled_oracle_error(w, out, nblocks=16, seed=0, stack=None):
    """Oracle check of a large stack on a seeded sample of C blocks: the entries that update those blocks
    are run through the double-double oracle, and the sampled blocks of out are compared (max normalised
    error over them)."""
    stack = w.stack if stack is None else stack
    cblk = (np.asarray(stack)[:, 2].astype(np.int64) - 1) // (w.m * w.n)
    present = np.unique(cblk)
    pick = np.random.default_rng(seed).choice(present, size=min(nblocks, present.size), replace=False)
    sub = np.ascontiguousarray(np.asarray(stack)[np.isin(cblk, pick)])
    hi, lo = oracle.stack_dd(w.m, w.n, w.k, sub, w.a, w.b, c_in=w.c.ravel())
    mn = w.m * w.n
    idx = (pick[:, None] * mn + np.arange(mn)[None, :]).ravel()
    return oracle.error(np.ascontiguousarray(out.ravel()[idx]), hi.ravel()[idx], lo.ravel()[idx])
