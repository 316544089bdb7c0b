"""ORACLE: test infrastructure only; never the product path.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this package. It wraps
liboracle.so (stack_oracle.c, ozaki_oracle.c, mlp_oracle.c, cpu_baseline.c): a double-double CPU
restatement of the reference's stack multiply (samples/smm/acc_bench.c:357-418 host path,
kernels/multiply.cl:12-20 layout), its generators (acc_bench.h:30-79), a double-double GEMM, Ozaki scheme I restated
(samples/ozaki/kernels/ozaki1_int8.cl, ozaki_common.cl:789-825), the predictor network's loss and
gradient in double precision, and the timed CPU baseline.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

PLANT_NONE, PLANT_BLAYOUT, PLANT_DROPK, PLANT_BASE0, PLANT_NOACC = 0, 1, 2, 3, 4

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-B"], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i, l, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_long, ctypes.c_double
        L.oracle_stack_dd.argtypes = [i, i, i, i, i, vp, vp, vp, i, vp, l, vp, vp, i]
        L.oracle_stack_dd.restype = i
        L.oracle_error.argtypes = [i, l, vp, vp, vp, vp]
        L.oracle_error.restype = i
        L.oracle_transpose.argtypes = [i, i, i, vp, i, i, vp]
        L.oracle_glibc_rand.argtypes = [ctypes.c_uint, i, vp]
        L.oracle_init_stack.argtypes = [vp, i, i, vp, i, i, i, i, i, i]
        L.oracle_init_stack.restype = i
        L.oracle_init_mat.argtypes = [i, d, vp, i, i, d]
        L.oracle_cpu_stack.argtypes = [i, i, i, i, i, vp, vp, vp, vp, i, i, i]
        L.oracle_cpu_stack.restype = d
        L.oracle_max_threads.restype = i
        L.oracle_cpu_isa.restype = i
        L.oracle_gemm_dd.argtypes = [ctypes.c_char, ctypes.c_char, i, i, i, d, vp, i, vp, i, d, vp, i, vp, vp, i, vp, vp]
        L.oracle_gemm_dd.restype = i
        L.oracle_oz_digits.argtypes = [i, vp, i, vp, vp]
        L.oracle_oz_digits.restype = i
        L.oracle_oz_digits_ref.argtypes = [i, vp, i, vp, vp]
        L.oracle_oz_digits_ref.restype = i
        L.oracle_ozaki_gemm.argtypes = [i, ctypes.c_char, ctypes.c_char, i, i, i, d, vp, i, vp, i, d, vp, i, i, i, i, i]
        L.oracle_ozaki_gemm.restype = i
        L.oracle_ozaki2_gemm.argtypes = [i, ctypes.c_char, ctypes.c_char, i, i, i, d, vp, i, vp, i, d, vp, i, i, i, i]
        L.oracle_ozaki2_gemm.restype = i
        L.oracle_mlp_loss_grad.argtypes = [i, vp, vp, vp, vp, vp, vp, i, i, vp, vp, i]
        L.oracle_mlp_loss_grad.restype = i
        _lib = L
    return _lib


def _tid(dtype):
    return 3 if np.dtype(dtype) == np.float64 else 1


def stack_dd(m, n, k, stack, a, b, c_in=None, b_rowmajor=False, plant=PLANT_NONE):
    """Double-double C after applying the stack: returns (hi, lo) float64 arrays shaped like c_in.

    a: A pool (column-major blocks), b: B pool (column-major k-by-n unless b_rowmajor),
    c_in: initial C pool (defaults to zeros of the size the stack addresses).
    """
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    assert a.dtype == b.dtype
    stack = np.ascontiguousarray(stack, dtype=np.int32)
    if c_in is None:
        nelems = int(stack[:, 2].max()) - 1 + m * n if len(stack) else m * n
        c_in = np.zeros(nelems, dtype=a.dtype)
    c_in = np.ascontiguousarray(c_in, dtype=a.dtype)
    hi = np.empty(c_in.size, dtype=np.float64)
    lo = np.empty(c_in.size, dtype=np.float64)
    rc = lib().oracle_stack_dd(_tid(a.dtype), m, n, k, len(stack), stack.ctypes.data, a.ctypes.data, b.ctypes.data,
                               1 if b_rowmajor else 0, c_in.ctypes.data, c_in.size, hi.ctypes.data, lo.ctypes.data,
                               plant)
    assert rc == 0
    return hi.reshape(c_in.shape), lo.reshape(c_in.shape)


def error(test, hi, lo):
    """(max normalised error, normwise error, max|ref|, max abs error) of test against hi+lo."""
    test = np.ascontiguousarray(test)
    out = np.zeros(4, dtype=np.float64)
    lib().oracle_error(_tid(test.dtype), test.size, test.ctypes.data, np.ascontiguousarray(hi).ctypes.data,
                       np.ascontiguousarray(lo).ctypes.data, out.ctypes.data)
    return tuple(out)


def transpose(m, n, trs, offset, data):
    trs = np.ascontiguousarray(trs, dtype=np.int32)
    assert data.flags.c_contiguous
    rc = lib().oracle_transpose(_tid(data.dtype), m, n, trs.ctypes.data, offset, len(trs) - offset, data.ctypes.data)
    assert rc == 0
    return data


def glibc_rand(seed, count):
    out = np.empty(count, dtype=np.int32)
    lib().oracle_glibc_rand(seed, count, out.ctypes.data)
    return out


def init_stack(stack_size, rnd, mn, mk, kn, nc, na, nb):
    rnd = np.ascontiguousarray(rnd, dtype=np.int32)
    out = np.empty((stack_size, 3), dtype=np.int32)
    rc = lib().oracle_init_stack(out.ctypes.data, stack_size, len(rnd), rnd.ctypes.data, mn, mk, kn, nc, na, nb)
    assert rc == 0
    return out


def init_mat(dtype, seed, m, n, scale):
    out = np.empty(m * n, dtype=dtype)
    lib().oracle_init_mat(_tid(dtype), float(seed), out.ctypes.data, m, n, float(scale))
    return out


def cpu_stack(m, n, k, stack, a, b, c, nrepeat=1, warmup=0, nthreads=0):
    """Reference host path (B column-major k-by-n) applied in place to c; returns timed seconds."""
    stack = np.ascontiguousarray(stack, dtype=np.int32)
    assert a.flags.c_contiguous and b.flags.c_contiguous and c.flags.c_contiguous
    return lib().oracle_cpu_stack(_tid(a.dtype), m, n, k, len(stack), stack.ctypes.data, a.ctypes.data,
                                  b.ctypes.data, c.ctypes.data, nrepeat, warmup, nthreads)


def max_threads():
    return lib().oracle_max_threads()


def cpu_isa():
    """instruction set of the CPU baseline's entry kernel on this host"""
    return ("x86-64", "avx2+fma", "avx512+fma")[lib().oracle_cpu_isa()]


def gemm_dd(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, rows=None, cols=None):
    """Double-double reference of C = alpha*op(A)*op(B) + beta*C (column-major); (hi, lo) for all
    entries (shape n-by-m... returned column-major as flat arrays) or for the listed (rows, cols)."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    c = None if c is None else np.ascontiguousarray(c, dtype=np.float64)
    if rows is not None:
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        count = len(rows)
    else:
        count = m * n
    hi = np.empty(count, dtype=np.float64)
    lo = np.empty(count, dtype=np.float64)
    rc = lib().oracle_gemm_dd(transa.encode(), transb.encode(), m, n, k, alpha, a.ctypes.data, lda, b.ctypes.data, ldb,
                              beta, None if c is None else c.ctypes.data, ldc, hi.ctypes.data, lo.ctypes.data,
                              0 if rows is None else count, None if rows is None else rows.ctypes.data,
                              None if rows is None else cols.ctypes.data)
    assert rc == 0
    return hi, lo


OZ_PLANT_NONE, OZ_PLANT_DROPLEVEL, OZ_PLANT_WEIGHT, OZ_PLANT_ROUND, OZ_PLANT_TRANSB = 0, 1, 2, 3, 4


def oz_digits(v, s):
    """(exponent, digits[s][n]) of one row in the device formulation (x = v/2^ex, d = trunc(128x))."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty((s, v.size), dtype=np.int8)
    ex = ctypes.c_int(0)
    assert lib().oracle_oz_digits(v.size, v.ctypes.data, s, out.ctypes.data, ctypes.byref(ex)) == 0
    return ex.value, out


def oz_digits_ref(v, s):
    """(max biased exponent, digits[s][n]) of one row in the reference's aligned-mantissa formulation."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.empty((s, v.size), dtype=np.int8)
    ex = ctypes.c_int(0)
    assert lib().oracle_oz_digits_ref(v.size, v.ctypes.data, s, out.ctypes.data, ctypes.byref(ex)) == 0
    return ex.value, out


def ozaki_gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s, trim=0, maxk=0, plant=OZ_PLANT_NONE):
    """Ozaki scheme I emulated with the device's operation order; returns the updated copy of c
    (column-major flat, float64 or float32 like a)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    out = np.array(c, dtype=a.dtype, copy=True, order="C").reshape(-1)
    use_double = 1 if a.dtype == np.float64 else 0
    rc = lib().oracle_ozaki_gemm(use_double, transa.encode(), transb.encode(), m, n, k, alpha, a.ctypes.data, lda,
                                 b.ctypes.data, ldb, beta, out.ctypes.data, ldc, s, trim, maxk, plant)
    assert rc == 0
    return out


def ozaki_stack(m, n, k, stack, a, b, c, s, trim=0, dtype=np.float64):
    """Scheme I over a parameter stack: every entry (1-based element offsets a, b, c) in stack order is
    the scheme I product of its A block (m x k column-major) and B block (k x n row-major, i.e. an n x k
    column-major store read transposed) added to its C block, C = fma(1, v, 1 C) as ozaki_gemm computes
    with alpha = beta = 1. Returns the updated copy of the flat C pool."""
    a = np.ascontiguousarray(a, dtype=dtype).reshape(-1)
    b = np.ascontiguousarray(b, dtype=dtype).reshape(-1)
    out = np.array(c, dtype=dtype, copy=True).reshape(-1)
    for ao, bo, co in np.asarray(stack, dtype=np.int64).reshape(-1, 3):
        ab = a[ao - 1:ao - 1 + m * k]
        bb = b[bo - 1:bo - 1 + k * n]
        out[co - 1:co - 1 + m * n] = ozaki_gemm("N", "T", m, n, k, 1.0, ab, m, bb, n, 1.0, out[co - 1:co - 1 + m * n],
                                                m, s, trim=trim)
    return out


OZ2_PLANT_NONE, OZ2_PLANT_ROUND, OZ2_PLANT_SIGNB, OZ2_PLANT_SCALE, OZ2_PLANT_FPSUM = 0, 1, 2, 3, 4


def ozaki2_gemm(transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, trim=0, kpass=0, plant=OZ2_PLANT_NONE):
    """Ozaki scheme II (CRT): the exact integer product of the aligned mantissas, rounded once, per K pass
    of kpass (0: all of K); returns the updated copy of c (column-major flat, float64 or float32 like a)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    out = np.array(c, dtype=a.dtype, copy=True, order="C").reshape(-1)
    use_double = 1 if a.dtype == np.float64 else 0
    rc = lib().oracle_ozaki2_gemm(use_double, transa.encode(), transb.encode(), m, n, k, alpha, a.ctypes.data, lda,
                                  b.ctypes.data, ldb, beta, out.ctypes.data, ldc, trim, kpass, plant)
    assert rc == 0
    return out


MLP_PLANT_NONE, MLP_PLANT_NORELU, MLP_PLANT_NOBIAS = 0, 1, 2


def mlp_loss_grad(width, params, x, tgt, seg, hw, plant=MLP_PLANT_NONE):
    """(mean loss, gradient) of the predictor network in double precision."""
    width = np.ascontiguousarray(width, dtype=np.int32)
    params = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    tgt = np.ascontiguousarray(tgt, dtype=np.float64)
    seg = np.ascontiguousarray(seg, dtype=np.int32)
    hw = np.ascontiguousarray(hw, dtype=np.float64)
    grad = np.zeros_like(params)
    loss = ctypes.c_double(0)
    rc = lib().oracle_mlp_loss_grad(len(width) - 1, width.ctypes.data, params.ctypes.data, x.ctypes.data,
                                    tgt.ctypes.data, seg.ctypes.data, hw.ctypes.data, len(seg) - 1, x.shape[0],
                                    grad.ctypes.data, ctypes.byref(loss), plant)
    assert rc == 0
    return loss.value, grad
