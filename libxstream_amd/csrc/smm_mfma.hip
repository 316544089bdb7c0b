/*
 * DBCSR stack multiply on gfx950 matrix cores: C[c_i] += A[a_i] * B[b_i] with MFMA 16x16x4
 * (v_mfma_f64_16x16x4f64 for FP64, v_mfma_f32_16x16x4f32 for FP32).
 *
 * Same semantics as smm_stack.hip (samples/smm/kernels/multiply.cl:12-15: A m-by-k column-major,
 * B k-by-n row-major, C m-by-n column-major, 1-based element offsets); a different design:
 *   - a wavefront covers the C block with 16x16 output tiles (TI x TJ tiles, padded); a
 *     work-group of NWI wavefronts splits the tile rows, each wavefront walking the same entries
 *     independently (no LDS, no barriers);
 *   - MFMA operands are loaded straight from global memory into registers: an A fragment is 16
 *     consecutive rows of 4 columns of the column-major A block, a B fragment 16 consecutive
 *     columns of 4 rows of the row-major B block, so every load is a set of 128-byte rows;
 *   - loads roll: as soon as a k-step's fragments have been fed to the matrix core, the same
 *     registers are refilled with the next unit of work (next k-stage, or the next stack entry),
 *     so one entry's loads are in flight while the previous entry is multiplied;
 *   - stack indices are wave-uniform scalar loads, fetched one entry ahead;
 *   - C-run ownership as in smm_stack.hip: a work-group owns the runs whose first entry lies in
 *     [w*bs, (w+1)*bs), sums the run in registers (per-entry products, segments of 16 entries)
 *     and flushes it once.
 *
 * Output layout of one 16x16 tile, lane l (0..63), register r (0..3), column j = l % 16:
 *   FP64 row = l/16 + 4r, FP32 row = 4(l/16) + r (checked on gfx950 with a layout probe).
 */
#include <hip/hip_runtime.h>
#include "lxs_kernels.h"

#if !defined(LXS_REAL)
#  define LXS_REAL double
#  define LXS_TAG f64
#endif
#define LXS_CAT2(A, B) A##B
#define LXS_CAT(A, B) LXS_CAT2(A, B)

namespace {

typedef LXS_REAL real_t;
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

struct mfma_entry_t {
  int ti, tj, nwi, kb, m, n, k; /* m = n = k = 0: runtime shape with m <= 16*ti, n <= 16*tj */
  const void* fn;
};

template <typename T>
struct mfma;
template <>
struct mfma<double> {
  typedef d4_t acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct mfma<float> {
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

template <typename T, int TI, int TJ, int NWI, int KB, int MC, int NC, int KC>
__global__ void __launch_bounds__(64 * NWI) smm_mfma(T* __restrict__ cdata, const T* __restrict__ adata,
  const T* __restrict__ bdata, const int* __restrict__ stack, int stack_size, int bs, int m_rt, int n_rt, int k_rt,
  int kt_unused, int atomic) {
  typedef typename mfma<T>::acc_t acc_t;
  constexpr int RI = TI / NWI;               /* tile rows per wavefront */
  constexpr bool SEG = (RI * TJ <= 4);       /* segment level of the run sum (registers permitting) */
  constexpr int SEG_LEN = 16;
  const int M = (0 < MC ? MC : m_rt), N = (0 < NC ? NC : n_rt), K = (0 < KC ? KC : k_rt);
  const int KS = (K + 3) >> 2;               /* k-steps of 4 */
  const int NS = (KS + KB - 1) / KB;         /* k-stages per entry */
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 15, lk = lane >> 4;
  const int ti0 = wave * RI;
  (void)kt_unused;

  const int begin = blockIdx.x * bs;
  const int end = min(begin + bs, stack_size);
  /* first run head in [begin, end) */
  int h = end;
  for (int base = begin; base < end; base += 64) {
    const int i = base + lane;
    bool head = false;
    if (i < end) head = (0 == i) || (stack[3 * i + 2] != stack[3 * i - 1]);
    const unsigned long long mask = __ballot(head);
    if (0 != mask) {
      h = base + __ffsll((long long)mask) - 1;
      break;
    }
  }
  if (h >= end) return;

  acc_t acc[RI][TJ], seg[RI][TJ], prod[RI][TJ];
  T fa[KB][RI], fb[KB][TJ];
  const acc_t zero = {T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int ri = 0; ri < RI; ++ri) {
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) acc[ri][tj] = seg[ri][tj] = prod[ri][tj] = zero;
  }

  /* fragments of k-step ks of the entry at offsets (a0, b0) into slot kb */
#define LXS_LOAD_KSTEP(KBI, A0, B0, KSTEP) \
  do { \
    const int k_ = 4 * (KSTEP) + lk; \
    _Pragma("unroll") for (int ri = 0; ri < RI; ++ri) { \
      const int i_ = 16 * (ti0 + ri) + li; \
      fa[KBI][ri] = (k_ < K && i_ < M) ? adata[(A0) + k_ * M + i_] : T(0); \
    } \
    _Pragma("unroll") for (int tj = 0; tj < TJ; ++tj) { \
      const int j_ = 16 * tj + li; \
      fb[KBI][tj] = (k_ < K && j_ < N) ? bdata[(B0) + k_ * N + j_] : T(0); \
    } \
  } while (0)

  int e = h;
  int a0 = stack[3 * e + 0] - 1, b0 = stack[3 * e + 1] - 1, c0 = stack[3 * e + 2];
  bool have1 = (e + 1 < stack_size);
  int a1 = 0, b1 = 0, c1 = -1;
  if (have1) {
    a1 = stack[3 * e + 3] - 1;
    b1 = stack[3 * e + 4] - 1;
    c1 = stack[3 * e + 5];
  }
  int s = 0, nent = 0;
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    if (kb < KS) LXS_LOAD_KSTEP(kb, a0, b0, kb);
  }

  for (;;) {
    const bool same_entry = (s + 1 < NS);
    const bool run_ends = !same_entry && (!have1 || c1 != c0);
    const bool own_next = have1 && (c1 == c0 || e + 1 < end);
    const bool has_next = same_entry || own_next;
    const int na = (same_entry ? a0 : a1), nb = (same_entry ? b0 : b1), ns = (same_entry ? s + 1 : 0);
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
      const int ks = s * KB + kb;
      if (ks < KS) {
#pragma unroll
        for (int ri = 0; ri < RI; ++ri) {
#pragma unroll
          for (int tj = 0; tj < TJ; ++tj) prod[ri][tj] = mfma<T>::mma(fa[kb][ri], fb[kb][tj], prod[ri][tj]);
        }
      }
      /* refill this slot with the next unit's fragments while the matrix core works */
      if (has_next && ns * KB + kb < KS) LXS_LOAD_KSTEP(kb, na, nb, ns * KB + kb);
    }
    if (same_entry) {
      ++s;
      continue;
    }
    /* entry done: fold its product into the run sum */
#pragma unroll
    for (int ri = 0; ri < RI; ++ri) {
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj) {
        if (SEG) {
          seg[ri][tj] += prod[ri][tj];
        }
        else {
          acc[ri][tj] += prod[ri][tj];
        }
        prod[ri][tj] = zero;
      }
    }
    if (SEG && (0 == (++nent % SEG_LEN) || run_ends)) {
#pragma unroll
      for (int ri = 0; ri < RI; ++ri) {
#pragma unroll
        for (int tj = 0; tj < TJ; ++tj) {
          acc[ri][tj] += seg[ri][tj];
          seg[ri][tj] = zero;
        }
      }
    }
    if (run_ends) {
      T* const cblk = cdata + (c0 - 1);
#pragma unroll
      for (int ri = 0; ri < RI; ++ri) {
#pragma unroll
        for (int tj = 0; tj < TJ; ++tj) {
          const int j = 16 * tj + li;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 16 * (ti0 + ri) + mfma<T>::row(lane, r);
            if (i < M && j < N) {
              T* const p = cblk + j * M + i;
              if (atomic) {
                unsafeAtomicAdd(p, acc[ri][tj][r]);
              }
              else {
                *p += acc[ri][tj][r];
              }
            }
          }
          acc[ri][tj] = zero;
        }
      }
      nent = 0;
    }
    if (!own_next) break;
    ++e;
    a0 = a1;
    b0 = b1;
    c0 = c1;
    s = 0;
    have1 = (e + 1 < stack_size);
    if (have1) {
      a1 = stack[3 * e + 3] - 1;
      b1 = stack[3 * e + 4] - 1;
      c1 = stack[3 * e + 5];
    }
  }
#undef LXS_LOAD_KSTEP
}

#define LXS_MG(TI, TJ, NWI, KB) {TI, TJ, NWI, KB, 0, 0, 0, (const void*)&smm_mfma<real_t, TI, TJ, NWI, KB, 0, 0, 0>},
#define LXS_MS(M, N, K, TI, TJ, NWI, KB) {TI, TJ, NWI, KB, M, N, K, (const void*)&smm_mfma<real_t, TI, TJ, NWI, KB, M, N, K>},

const mfma_entry_t mfma_table[] = {
  /* runtime shape (m <= 16*ti, n <= 16*tj), 4 k-steps (16 of k) per stage */
  LXS_MG(1, 1, 1, 4) LXS_MG(1, 2, 1, 4) LXS_MG(2, 1, 1, 4) LXS_MG(2, 2, 1, 4) LXS_MG(2, 2, 2, 4)
  LXS_MG(3, 3, 1, 2) LXS_MG(3, 3, 3, 4) LXS_MG(4, 4, 4, 4) LXS_MG(4, 4, 2, 2) LXS_MG(5, 5, 5, 4)
  /* compile-time shapes of the stack workloads (whole k per stage where registers allow) */
  LXS_MS(5, 5, 5, 1, 1, 1, 2) LXS_MS(13, 13, 13, 1, 1, 1, 4)
  LXS_MS(23, 23, 23, 2, 2, 1, 6) LXS_MS(23, 23, 23, 2, 2, 2, 6) LXS_MS(23, 23, 23, 2, 2, 1, 3)
  LXS_MS(32, 32, 32, 2, 2, 1, 8) LXS_MS(32, 32, 32, 2, 2, 2, 8) LXS_MS(32, 32, 32, 2, 2, 1, 4)
  LXS_MS(64, 64, 64, 4, 4, 4, 4) LXS_MS(64, 64, 64, 4, 4, 2, 4)
  LXS_MS(4, 9, 22, 1, 1, 1, 6) LXS_MS(45, 45, 13, 3, 3, 1, 4) LXS_MS(45, 45, 13, 3, 3, 3, 4)
};

} /* namespace */

extern "C" const void* LXS_CAT(lxs_mfma_table_, LXS_TAG)(int* count) {
  *count = (int)(sizeof(mfma_table) / sizeof(*mfma_table));
  return mfma_table;
}
