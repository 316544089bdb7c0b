/*
 * Ozaki scheme I: FP64/FP32 GEMM emulated with int8 matrix-core products (MI355X, gfx950).
 *
 * Drop-in for the reference's samples/ozaki/ozaki_opencl.h:328-440 (same entry points, argument
 * meaning and return convention: EXIT_SUCCESS / EXIT_FAILURE). The context is caller-allocated as in
 * the reference; its public fields carry the resolved settings, `impl` owns the device state.
 *
 * Scheme: every row of op(A) and column of op(B) is scaled by a power of two and cut into nslices
 * signed 7-bit digits (int8 planes); slice pairs (p, q) with p + q <= cutoff = 2(nslices-1) - oztrim
 * are multiplied with int8 MFMA, summed exactly per level p + q and recombined in FP64.
 * Scheme 2 (kind 2) turns the operands into int8 residues modulo 16 pairwise coprime moduli <= 256 and
 * reconstructs the exact integer product by CRT before one rounding to FP64; kind 0/3 picks a scheme per
 * call (samples/ozaki/ozaki_gemm.c:520-587).
 */
#ifndef OZAKI_H
#define OZAKI_H

#include "libxstream.h"
#include <stddef.h>

/** samples/ozaki/ozaki_opencl.h:58 */
typedef enum ozaki_flags_t { OZAKI_TRIANGULAR = 1, OZAKI_SYMMETRIZE = 2 } ozaki_flags_t;

/**
 * Number of slice products for nslices and cutoff co under flags (samples/ozaki/ozaki_opencl.h:60-73):
 * flags 0 and 3 both cover every pair with p + q <= co.
 */
static inline int ozaki_count_pairs(int nslices, int co, int flags) {
  int p, q, count = 0;
  for (p = 0; p < nslices && p <= co; ++p) {
    for (q = ((0 != (flags & OZAKI_TRIANGULAR)) ? p : 0); q < nslices && p + q <= co; ++q) {
      count += ((0 != (flags & OZAKI_SYMMETRIZE)) && p != q) ? 2 : 1;
    }
  }
  return count;
}

typedef struct ozaki_context_t {
  int use_double; /* 1: fp64, 0: fp32 */
  int kind;       /* resolved scheme: 1 = mantissa slicing */
  int ndecomp;    /* number of decomposition components (= nslices) */
  int nslices;    /* int8 digit planes per operand, 1..8 (default fp64: 8, fp32: 4) */
  int nprimes;    /* CRT primes: 0 (Scheme 2 is not built) */
  int ozflags;    /* OZAKI_TRIANGULAR | OZAKI_SYMMETRIZE: 0 and 3 compute the full sum */
  int oztrim;     /* levels dropped from the top of the pair triangle, 0..2(nslices-1) */
  int verbosity;  /* 0: quiet, 1: info, 2+: debug */
  int tm, tn;     /* output tile of one work-group (256 x 256) */
  int maxk;       /* K per preprocessing pass (0: all of K) */
  int flush;      /* k-blocks of 64 per int8 work item, 1..2080 (env LIBXSTREAM_OZAKI_FLUSH); results
                     do not depend on it: level sums are exact */
  void* impl;     /* device state owned by the context */
} ozaki_context_t;

/** samples/ozaki/ozaki_opencl.h:328; tm/tn: 0 or 256; ozgroups is accepted and ignored (Scheme 2) */
LIBXSTREAM_API int ozaki_init(ozaki_context_t* ctx, int tm, int tn, int use_double, int kind, int verbosity, int ndecomp,
  int ozflags, int oztrim, int ozgroups, int maxk);
/** samples/ozaki/ozaki_opencl.h:330 */
LIBXSTREAM_API void ozaki_destroy(ozaki_context_t* ctx);

/** samples/ozaki/ozaki_opencl.h:347: device bytes one call of this shape can need (upper bound) */
LIBXSTREAM_API size_t ozaki_scratch_size(const ozaki_context_t* ctx, char transa, char transb, int M, int N, int K,
  int lda, int ldb, int ldc);
/** samples/ozaki/ozaki_opencl.h:349: install caller-owned device memory (NULL reverts to the internal arena) */
LIBXSTREAM_API int ozaki_scratch_set(ozaki_context_t* ctx, void* dev_mem, size_t nbytes);
/** samples/ozaki/ozaki_opencl.h:358 */
LIBXSTREAM_API void* ozaki_scratch_get(const ozaki_context_t* ctx, size_t* nbytes);

/** samples/ozaki/ozaki_opencl.h:361-363 */
typedef struct ozaki_tile_t {
  int m, n;
} ozaki_tile_t;
/** samples/ozaki/ozaki_opencl.h:384: the output tile the GEMM kernel uses (fixed 256 x 256) */
LIBXSTREAM_API ozaki_tile_t ozaki_tile_select(const ozaki_context_t* ctx, int M, int N, int rtm, int rtn);
/** samples/ozaki/ozaki_opencl.h:411: N-panel width; the device path is not panelled, so N */
LIBXSTREAM_API int ozaki_npanel(const ozaki_context_t* ctx, int M, int N, int tm, int tn);

/**
 * samples/ozaki/ozaki_opencl.h:420: C = alpha op(A) op(B) + beta C, column-major, op = 'N'/'T'/'C'.
 * dev != 0: a, b, c are device pointers; dev == 0: host pointers (copied in and out on stream).
 * Enqueued on stream (NULL: the calling thread's default stream); the caller synchronizes.
 */
LIBXSTREAM_API int ozaki_gemm(ozaki_context_t* ctx, libxstream_stream_t* stream, char transa, char transb, int M, int N,
  int K, double alpha, const void* a, int lda, const void* b, int ldb, double beta, void* c, int ldc, int dev);

/**
 * No counterpart in the reference: ozaki_gemm for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc). Same arguments, scheme, kernels and
 * results as ozaki_gemm with dev = 1 (a, b, c device pointers); its scratch (digit or residue planes,
 * exponents, partial products) lives in a static device array of 64 MiB per device that the library
 * holds until the process ends, so the call allocates nothing. A call whose scratch would exceed 64 MiB
 * (FP64 Scheme 2 with 16 moduli: up to about M = N = K = 1024) fails with EXIT_FAILURE and a message.
 * Every call on one device shares that device's array: calls that overlap on one device (several streams
 * or threads) race and must be ordered by the caller.
 */
LIBXSTREAM_API int ozaki_gemm_noalloc(ozaki_context_t* ctx, libxstream_stream_t* stream, char transa, char transb,
  int M, int N, int K, double alpha, const void* a, int lda, const void* b, int ldb, double beta, void* c, int ldc);

/** samples/ozaki/ozaki_opencl.h:430 (3M method); alpha/beta point to (re, im); device pointers */
LIBXSTREAM_API int ozaki_gemm_complex(ozaki_context_t* ctx, libxstream_stream_t* stream, char transa, char transb, int M,
  int N, int K, const double* alpha, const void* a, int lda, const void* b, int ldb, const double* beta, void* c, int ldc);

/** samples/ozaki/ozaki_opencl.h:440: no preprocessing cache is kept, so nothing can go stale */
LIBXSTREAM_API void ozaki_invalidate_cache(ozaki_context_t* ctx, const void* a, const void* b);

/**
 * No counterpart in the reference (its comparator is the vendor BLAS): native FP64 matrix-core GEMM
 * (v_mfma_f64_16x16x4f64), C = alpha op(A) op(B) + beta C on device pointers, column-major.
 */
LIBXSTREAM_API int ozaki_dgemm_native(libxstream_stream_t* stream, char transa, char transb, int M, int N, int K,
  double alpha, const double* a, int lda, const double* b, int ldb, double beta, double* c, int ldc);

/**
 * No counterpart in the reference (its Ozaki sample multiplies whole matrices): Ozaki scheme I over a
 * DBCSR parameter stack, C[c] += A[a] B[b] for every entry, with the semantics of libsmm_acc_process
 * (include/acc_libsmm.h): stack of stack_size triplets of 1-based element offsets (a, b, c), A blocks
 * m x k column-major, B blocks in the kernel layout (k x n row-major, as libsmm_acc_transpose leaves
 * them), C blocks m x n column-major, 1 <= m, n, k <= 64, FP64 or FP32 as ctx->use_double. na and nb are
 * the blocks of the A and B pools (every block is sliced once per call; an entry whose A or B offset is
 * not a block start inside its pool contributes nothing). Each product is ozaki_gemm's scheme I with the
 * context's slice count and trim (8 slices over all pairs for a kind 2 context), added to C entry by
 * entry in stack order: C = C + v (FP32: rounded through FP64). A stack whose C offsets decrease
 * somewhere runs on one work-group, in order.
 */
LIBXSTREAM_API int ozaki_stack(ozaki_context_t* ctx, libxstream_stream_t* stream, const int* dev_stack, int stack_size,
  int m, int n, int k, const void* dev_a, int na, const void* dev_b, int nb, void* dev_c);

#endif /*OZAKI_H*/
