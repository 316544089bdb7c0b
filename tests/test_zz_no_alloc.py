"""GPU check of ozaki_gemm_noalloc (include/ozaki.h) through the C ABI: the entry point that keeps its
scratch in the library's static device array instead of allocating.

Bar: its C equals ozaki_gemm's (dev = 1) bit for bit on the same context and inputs, on two calls in a
row where the second is smaller (the array keeps the first call's contents), for Scheme 2 (residues) and
Scheme 1 (digit slices), with transposes and beta != 0. Device memory free before and after the calls is
unchanged, and a call that needs more than the array is refused with an error.
"""
import ctypes

import numpy as np
import pytest

import libxstream_amd as lx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def stream():
    lx.init()
    assert lx.device_count() > 0, "no HIP device visible"
    s = lx.Stream("noalloc")
    yield s
    s.close()
    lx.finalize()


def _operands(rng, ta, tb, m, n, k, e=10):
    """column-major storage of op(A) (m x k) and op(B) (k x n): mantissas [1, 2), row/column exponents in
    [-e, e], random signs; C uniform [-1, 1]"""
    A = rng.uniform(1, 2, (m, k)) * np.exp2(rng.integers(-e, e + 1, (m, 1))) * rng.choice([-1.0, 1.0], (m, k))
    B = rng.uniform(1, 2, (k, n)) * np.exp2(rng.integers(-e, e + 1, (1, n))) * rng.choice([-1.0, 1.0], (k, n))
    As = A.T if ta == "T" else A
    Bs = B.T if tb == "T" else B
    C = rng.uniform(-1, 1, (m, n))
    return (np.asfortranarray(As).ravel(order="F"), As.shape[0], np.asfortranarray(Bs).ravel(order="F"), Bs.shape[0],
            np.asfortranarray(C).ravel(order="F"))


def _free_bytes():
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    lx.check(lx.lib().libxstream_mem_info(ctypes.byref(free), ctypes.byref(total)), "libxstream_mem_info")
    return free.value


def _both(oz, stream, ta, tb, m, n, k, alpha, beta, seed):
    a, lda, b, ldb, c = _operands(np.random.default_rng(seed), ta, tb, m, n, k)
    da, db = lx.DeviceBuffer.from_array(a), lx.DeviceBuffer.from_array(b)
    dc1, dc2 = lx.DeviceBuffer.from_array(c), lx.DeviceBuffer.from_array(c)
    try:
        oz.gemm(ta, tb, m, n, k, alpha, da, lda, db, ldb, beta, dc1, m, stream=stream)
        stream.sync()
        free0 = _free_bytes()
        oz.gemm_noalloc(ta, tb, m, n, k, alpha, da, lda, db, ldb, beta, dc2, m, stream=stream)
        stream.sync()
        free1 = _free_bytes()
        ref, out = np.empty_like(c), np.empty_like(c)
        dc1.download(ref)
        dc2.download(out)
    finally:
        for x in (da, db, dc1, dc2):
            x.free()
    return ref, out, free0 - free1


@pytest.mark.parametrize("kind", [2, 1], ids=["scheme2", "scheme1"])
def test_noalloc_matches_ozaki_gemm(stream, kind):
    oz = lx.Ozaki(kind=kind, nslices=7 if kind == 1 else 0)
    try:
        # first call near the array's size for the scheme, then a smaller one on the array's stale contents
        first = (1000, 1024, 1000) if kind == 2 else (480, 500, 512)
        for (m, n, k), ta, tb, alpha, beta, seed in ((first, "N", "N", 1.0, 0.0, 1),
                                                     ((200, 130, 90), "T", "N", -0.75, 0.5, 2),
                                                     ((64, 33, 300), "N", "T", 2.0, 1.0, 3)):
            ref, out, drop = _both(oz, stream, ta, tb, m, n, k, alpha, beta, seed)
            assert np.isfinite(ref).all()
            assert np.array_equal(ref.view(np.uint64), out.view(np.uint64)), (kind, m, n, k, ta, tb)
            assert drop < (1 << 20), "device memory dropped by %d bytes during ozaki_gemm_noalloc" % drop
    finally:
        oz.close()


def test_noalloc_rejects_oversize(stream):
    oz = lx.Ozaki(kind=2)
    m = n = k = 1400
    bufs = [lx.DeviceBuffer(8 * m * k), lx.DeviceBuffer(8 * k * n), lx.DeviceBuffer(8 * m * n)]
    try:
        with pytest.raises(lx.LibxstreamError, match="static scratch"):
            oz.gemm_noalloc("N", "N", m, n, k, 1.0, bufs[0], m, bufs[1], k, 0.0, bufs[2], m, stream=stream)
        stream.sync()
    finally:
        for x in bufs:
            x.free()
        oz.close()
