/*
 * Ozaki scheme I over a DBCSR parameter stack (gfx950): C[c] += A[a] B[b] for every stack entry, the
 * products emulated with int8 matrix cores. No counterpart in the reference (its Ozaki sample,
 * samples/ozaki/, multiplies whole matrices); the stack semantics are those of libsmm_acc_process
 * (samples/smm/kernels/multiply.cl:12-20, 1-based element offsets in triplets, A column-major m x k,
 * B in the kernel layout k x n row-major, C column-major m x n) and the digits those of ozaki_gemm
 * (per-row exponent, 7-bit signed digits, levels recombined in FP64).
 *
 *   ozs_slice   every block of a pool -> s digit planes of 64 x 64 int8 (rows of A, columns of B;
 *               k contiguous) and the 64 row exponents; padding rows and k are zero.
 *   ozs_check   whether C offsets ever decrease (the stack is then not grouped by C).
 *   ozs_stack   persistent: work-group g owns the runs (entries of one C offset) that start in its equal
 *               share of the stack and walks each to its end; a run's C block stays in registers
 *               (FP64, or FP32 updated through FP64); per entry the 2s digit planes of its A and B
 *               blocks move into LDS by DMA (double-buffered across entries), every level sum
 *               S_L = sum over pairs p + q = L of A_p B_q^T (64 x 64 x 64) accumulates exactly in its own
 *               int32 tile (v_mfma_i32_32x32x32_i8, 4 wavefronts of 32 x 32), then
 *               cd = sum over L = cutoff .. 0 of 2^(-7(L+2)) S_L in FP64 and C += 2^(ea_i + eb_j) cd.
 *               A stack that is not sorted by C is processed by one work-group, run after run, in order.
 * Every element of C is therefore updated entry by entry in stack order, reproducibly.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "lxs_kernels.h"

namespace {

typedef int i4_t __attribute__((ext_vector_type(4)));
typedef int i16_t __attribute__((ext_vector_type(16)));

constexpr int SB = 64;           /* block rows, columns and k, padded */
constexpr int SPL = SB * SB;     /* bytes of one digit plane */
constexpr int SNONE = -(1 << 30);

__device__ __forceinline__ int ozs_expo(double v) { /* |v| < 2^ex for a normal finite v */
  const long long bits = __double_as_longlong(v);
  const int be = (int)((bits >> 52) & 0x7FF);
  return (0 == be || 0x7FF == be) ? SNONE : be - 1022;
}
__device__ __forceinline__ int ozs_expo(float v) {
  const int bits = __float_as_int(v);
  const int be = (bits >> 23) & 0xFF;
  return (0 == be || 0xFF == be) ? SNONE : be - 126;
}

/*
 * Pool slicing: block b of the pool, element (r, kk) at p[b * bsz + r * rs + kk * ks] for r < R, kk < K.
 * Thread t owns row t / 4 and k = 16 (t % 4) .. +15; the row exponent is the maximum over 4 lanes.
 * Planes at planes[((b * s + pl) * 64 + r) * 64 + kk]; exponents at ex[b * 64 + r].
 */
template <typename T>
__global__ void __launch_bounds__(256) ozs_slice(const T* __restrict__ p, long bsz, int rs, int ks, int R, int K, int s,
  int8_t* __restrict__ planes, int* __restrict__ ex) {
  const long blk = blockIdx.x;
  const int t = threadIdx.x, r = t >> 2, kq = (t & 3) * 16;
  const T* const src = p + blk * bsz;
  double v[16];
  int e = SNONE;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const T x = (r < R && kq + i < K) ? src[(long)r * rs + (long)(kq + i) * ks] : T(0);
    const int ev = ozs_expo(x);
    v[i] = (SNONE != ev ? (double)x : 0.0);
    e = max(e, ev);
  }
  e = max(e, __shfl_xor(e, 1));
  e = max(e, __shfl_xor(e, 2));
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = (SNONE != e ? ldexp(v[i], -e) : 0.0); /* |x| < 1, exact */
  for (int pl = 0; pl < s; ++pl) {
    union {
      int8_t b[16];
      i4_t w;
    } d;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const double y = v[i] * 128.0;
      const double q = trunc(y); /* |q| <= 127 */
      d.b[i] = (int8_t)q;
      v[i] = y - q; /* exact */
    }
    *(i4_t*)(planes + ((blk * s + pl) * SB + r) * SB + kq) = d.w;
  }
  if (0 == (t & 3)) ex[blk * SB + r] = e;
}

__global__ void ozs_clear(int* __restrict__ info) { info[threadIdx.x] = 0; }

/* info[0] = 1 when some C offset is smaller than its predecessor's */
__global__ void __launch_bounds__(256) ozs_check(const int* __restrict__ stack, int S, int* __restrict__ info) {
  int bad = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x + 1; i < S; i += gridDim.x * 256) {
    bad |= (stack[3 * i + 2] < stack[3 * (i - 1) + 2]) ? 1 : 0;
  }
  if (0 != __syncthreads_or(bad) && 0 == threadIdx.x) atomicOr(info, 1);
}

__device__ __forceinline__ void ozs_dma16(const int8_t* src, int8_t* lds_dst) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
    (__attribute__((address_space(3))) void*)lds_dst, 16, 0, 0);
}
__device__ __forceinline__ void ozs_dma4(const int* src, int8_t* lds_dst) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
    (__attribute__((address_space(3))) void*)lds_dst, 4, 0, 0);
}
constexpr int SEX = 512; /* LDS bytes per wavefront for an entry's exponent rows (A rows, then B columns) */

/*
 * LDS: two entry buffers of 2 NS planes (A planes, then B planes), 4 KB each, and each wavefront's copy of the
 * entry's exponent rows (2 x 256 bytes); row r of a plane holds its
 * 4 chunks of 16 bytes with chunk c at position c ^ ((r >> 2) & 3). Wave w moves rows 16 w .. 16 w + 15 of
 * every plane (one DMA instruction per plane: lane l writes position l % 4 of row 16 w + l / 4).
 */
template <typename T, int NS, int CUT>
__global__ void __launch_bounds__(256) ozs_stack(const int* __restrict__ stack, int S, const int* __restrict__ info,
  const int8_t* __restrict__ pa, const int* __restrict__ ea, const int8_t* __restrict__ pb, const int* __restrict__ eb,
  int cutoff_rt, int m, int n, int k, int na, int nb, T* __restrict__ c) {
  /* the level cutoff: compiled in (CUT >= 0), so the pair loop below has no conditional matrix-core work,
     or taken at run time (CUT < 0) */
  const int cutoff = (0 <= CUT ? CUT : cutoff_rt);
  __shared__ __attribute__((aligned(1024))) int8_t lds0[2 * NS * SPL + 4 * SEX], lds1[2 * NS * SPL + 4 * SEX];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wi = wave & 1, wj = wave >> 1;
  const int unsorted = info[0];
  const unsigned int mk = (unsigned int)(m * k), kn = (unsigned int)(k * n);
  /* this work-group's share of the stack (all of it for work-group 0 when the stack is not grouped by C) */
  const int share = (S + (int)gridDim.x - 1) / (int)gridDim.x;
  const int lo = (0 != unsorted ? 0 : (int)blockIdx.x * share);
  const int hi = (0 != unsorted ? S : (lo + share < S ? lo + share : S));
  if (0 != unsorted && 0 != blockIdx.x) return;
  /* this lane's element of the 32 x 32 tile: column i (A rows), rows j (B columns) per register */
  const int i = 32 * wi + (lane & 31);
  const int drow = 16 * wave + (lane >> 2), dchunk = (lane & 3) ^ ((lane >> 4) & 3);
  const int frow = lane & 31, fswz = (frow >> 2) & 3;
  const int fpos0 = (lane >> 5) ^ fswz, fpos1 = (2 + (lane >> 5)) ^ fswz;
  /* run boundaries found 64 entries at a time (one lane each, then a ballot) */
  int e0 = lo; /* the first run that starts in [lo, hi) */
  for (;;) {
    const int x = e0 + lane;
    const bool start = (x < hi) && (0 == x || stack[3 * x + 2] != stack[3 * (x - 1) + 2]);
    const unsigned long long bal = __ballot(start);
    if (0 != bal) {
      e0 += __ffsll((long long)bal) - 1;
      break;
    }
    e0 += 64;
    if (e0 >= hi) {
      e0 = hi;
      break;
    }
  }
  while (e0 < hi) { /* the run of entries e0 .. e1 - 1 (one C offset), walked to its end */
    const int cofs = stack[3 * e0 + 2];
    int e1 = e0 + 1;
    for (;;) {
      const int x = e1 + lane;
      const unsigned long long bal = __ballot(x >= S || stack[3 * x + 2] != cofs);
      if (0 != bal) {
        e1 += __ffsll((long long)bal) - 1;
        break;
      }
      e1 += 64;
    }
    const long cbase = (long)cofs - 1;
    double creg[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int j = 32 * wj + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      creg[r] = (i < m && j < n) ? (double)c[cbase + i + (long)j * m] : 0.0;
    }
    /* block indices of entry e, -1 when its offsets are not block starts inside the pools */
    auto blocks = [&](int e, int& ab, int& bb) {
      const int ao = __builtin_amdgcn_readfirstlane(stack[3 * e] - 1), bo = __builtin_amdgcn_readfirstlane(stack[3 * e + 1] - 1);
      const unsigned int aq = (unsigned int)ao / mk, bq = (unsigned int)bo / kn;
      ab = (0 <= ao && (unsigned int)ao == aq * mk && aq < (unsigned int)na) ? (int)aq : -1;
      bb = (0 <= bo && (unsigned int)bo == bq * kn && bq < (unsigned int)nb) ? (int)bq : -1;
    };
    /* no branch around the DMA (an entry without valid blocks loads block 0, unused), so the compiler
       can count the outstanding loads on every path */
    auto dma = [&](int8_t* buf, int ab, int bb) {
      ab = (0 <= ab ? ab : 0);
      bb = (0 <= bb ? bb : 0);
#pragma unroll
      for (int pl = 0; pl < NS; ++pl) {
        ozs_dma16(pa + ((long)ab * NS + pl) * SPL + drow * SB + 16 * dchunk, buf + pl * SPL + wave * 1024);
        ozs_dma16(pb + ((long)bb * NS + pl) * SPL + drow * SB + 16 * dchunk, buf + (NS + pl) * SPL + wave * 1024);
      }
      /* the 64 row exponents of the A block and column exponents of the B block, one copy per wavefront */
      ozs_dma4(ea + (long)ab * SB + lane, buf + 2 * NS * SPL + wave * SEX);
      ozs_dma4(eb + (long)bb * SB + lane, buf + 2 * NS * SPL + wave * SEX + SB * 4);
    };
    int ab, bb;
    blocks(e0, ab, bb);
    dma(lds0, ab, bb);
    for (int e = e0; e < e1; ++e) {
      int8_t* const cur = (0 == ((e - e0) & 1) ? lds0 : lds1);
      int8_t* const nxt = (0 == ((e - e0) & 1) ? lds1 : lds0);
      const int ac = ab, bc = bb;
      __builtin_amdgcn_s_waitcnt(0); /* own DMA of entry e, own reads of entry e - 1 */
      __syncthreads();              /* every wave's DMA of e has landed, every wave is done with e - 1 */
      /* entry e + 1 into the other buffer; the last entry of the run loads its own blocks again (unused).
         The loop has no other global loads, so nothing waits for this DMA before the next entry */
      blocks(e + 1 < e1 ? e + 1 : e, ab, bb);
      dma(nxt, ab, bb);
      if (0 <= ac && 0 <= bc) {
        /* exponents from LDS: row i of A (SNONE past m), columns 32 wj + 8 q + 4 (lane / 32) + 0..3 of B */
        const int* const lex = (const int*)(cur + 2 * NS * SPL + wave * SEX);
        const int exi = lex[i];
        i4_t exj4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) exj4[q] = *(const i4_t*)(lex + SB + 32 * wj + 8 * q + 4 * (lane >> 5));
        i4_t fa[NS][2], fb[NS][2];
#pragma unroll
        for (int pl = 0; pl < NS; ++pl) {
          const int8_t* const la = cur + pl * SPL + (32 * wi + frow) * SB;
          const int8_t* const lb = cur + (NS + pl) * SPL + (32 * wj + frow) * SB;
          fa[pl][0] = *(const i4_t*)(la + 16 * fpos0);
          fa[pl][1] = *(const i4_t*)(la + 16 * fpos1);
          fb[pl][0] = *(const i4_t*)(lb + 16 * fpos0);
          fb[pl][1] = *(const i4_t*)(lb + 16 * fpos1);
        }
        i16_t acc[2 * NS - 1];
#pragma unroll
        for (int L = 0; L < 2 * NS - 1; ++L) acc[L] = i16_t{0};
#pragma unroll
        for (int p = 0; p < NS; ++p) {
#pragma unroll
          for (int q = 0; q < NS; ++q) {
            if (p + q <= cutoff) {
              acc[p + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fb[q][0], fa[p][0], acc[p + q], 0, 0, 0);
              acc[p + q] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fb[q][1], fa[p][1], acc[p + q], 0, 0, 0);
            }
          }
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int j = 32 * wj + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          double cd = 0.0;
#pragma unroll
          for (int L = 2 * NS - 2; 0 <= L; --L) {
            if (L <= cutoff) cd = fma(ldexp(1.0, -7 * (L + 2)), (double)acc[L][r], cd);
          }
          const int exj = (j < n ? exj4[r >> 2][r & 3] : SNONE);
          const double v = (SNONE != exi && SNONE != exj) ? ldexp(cd, exi + exj) : 0.0;
          creg[r] = (double)(T)(creg[r] + v);
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int j = 32 * wj + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
      if (i < m && j < n) c[cbase + i + (long)j * m] = (T)creg[r];
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads(); /* the buffers of this run are free before the next run's first DMA */
    e0 = e1;
  }
}

/* NS slices with the level cutoff compiled in for the triangle (NS - 1) and all pairs (2 NS - 2) */
template <typename T, int NS>
void ozs_launch_ns(dim3 grid, hipStream_t st, const int* stack, int S, const int* info, const int8_t* pa,
  const int* ea, const int8_t* pb, const int* eb, int cutoff, int m, int n, int k, int na, int nb, T* c) {
  if (NS - 1 == cutoff) {
    ozs_stack<T, NS, NS - 1><<<grid, 256, 0, st>>>(stack, S, info, pa, ea, pb, eb, cutoff, m, n, k, na, nb, c);
  }
  else if (2 * (NS - 1) == cutoff) {
    ozs_stack<T, NS, 2 * (NS - 1)><<<grid, 256, 0, st>>>(stack, S, info, pa, ea, pb, eb, cutoff, m, n, k, na, nb, c);
  }
  else {
    ozs_stack<T, NS, -1><<<grid, 256, 0, st>>>(stack, S, info, pa, ea, pb, eb, cutoff, m, n, k, na, nb, c);
  }
}

template <typename T>
int ozs_launch_stack(int s, dim3 grid, hipStream_t st, const int* stack, int S, const int* info,
  const int8_t* pa, const int* ea, const int8_t* pb, const int* eb, int cutoff, int m, int n, int k, int na, int nb, T* c) {
  switch (s) {
    case 1: ozs_launch_ns<T, 1>(grid, st, stack, S, info, pa, ea, pb, eb, cutoff, m, n, k, na, nb, c); break;
    case 2: ozs_launch_ns<T, 2>(grid, st, stack, S, info, pa, ea, pb, eb, cutoff, m, n, k, na, nb, c); break;
    case 3: ozs_launch_ns<T, 3>(grid, st, stack, S, info, pa, ea, pb, eb, cutoff, m, n, k, na, nb, c); break;
    case 4: ozs_launch_ns<T, 4>(grid, st, stack, S, info, pa, ea, pb, eb, cutoff, m, n, k, na, nb, c); break;
    case 5: ozs_launch_ns<T, 5>(grid, st, stack, S, info, pa, ea, pb, eb, cutoff, m, n, k, na, nb, c); break;
    case 6: ozs_launch_ns<T, 6>(grid, st, stack, S, info, pa, ea, pb, eb, cutoff, m, n, k, na, nb, c); break;
    case 7: ozs_launch_ns<T, 7>(grid, st, stack, S, info, pa, ea, pb, eb, cutoff, m, n, k, na, nb, c); break;
    case 8: ozs_launch_ns<T, 8>(grid, st, stack, S, info, pa, ea, pb, eb, cutoff, m, n, k, na, nb, c); break;
    default: return LXS_ELIMIT;
  }
  return 0;
}

} /* namespace */

/*
 * A pool of na blocks m x k (column-major), B pool of nb blocks in the kernel layout (k x n row-major);
 * planes_a: na * s * 4 KB, planes_b: nb * s * 4 KB, ea: na * 64 ints, eb: nb * 64 ints, info: 1 int (all
 * device memory).
 */
extern "C" int lxs_ozs_process(int dtype, const int* stack, int S, int m, int n, int k, const void* a, int na,
  const void* b, int nb, void* c, int s, int cutoff, signed char* planes_a, int* ea, signed char* planes_b, int* eb,
  int* info, void* hip_stream) {
  const hipStream_t st = (hipStream_t)hip_stream;
  if (1 > m || SB < m || 1 > n || SB < n || 1 > k || SB < k || 1 > s || 8 < s || 0 > cutoff || 2 * (s - 1) < cutoff) {
    return LXS_ELIMIT;
  }
  if (0 >= S) return 0;
  if (0 >= na || 0 >= nb) return LXS_ELIMIT;
  const dim3 grid((unsigned int)(S < 1024 ? S : 1024));
  const unsigned int cgrid = (unsigned int)((S + 255) / 256 < 256 ? (S + 255) / 256 : 256);
  int result = 0;
  if (3 == dtype) { /* A rows: (r, kk) at r + kk m; B columns: (j, kk) at j + kk n */
    ozs_slice<double><<<na, 256, 0, st>>>((const double*)a, (long)m * k, 1, m, m, k, s, (int8_t*)planes_a, ea);
    ozs_slice<double><<<nb, 256, 0, st>>>((const double*)b, (long)k * n, 1, n, n, k, s, (int8_t*)planes_b, eb);
    ozs_clear<<<1, 1, 0, st>>>(info);
    ozs_check<<<cgrid, 256, 0, st>>>(stack, S, info);
    result = ozs_launch_stack<double>(s, grid, st, stack, S, info, (const int8_t*)planes_a, ea, (const int8_t*)planes_b,
      eb, cutoff, m, n, k, na, nb, (double*)c);
  }
  else if (1 == dtype) {
    ozs_slice<float><<<na, 256, 0, st>>>((const float*)a, (long)m * k, 1, m, m, k, s, (int8_t*)planes_a, ea);
    ozs_slice<float><<<nb, 256, 0, st>>>((const float*)b, (long)k * n, 1, n, n, k, s, (int8_t*)planes_b, eb);
    ozs_clear<<<1, 1, 0, st>>>(info);
    ozs_check<<<cgrid, 256, 0, st>>>(stack, S, info);
    result = ozs_launch_stack<float>(s, grid, st, stack, S, info, (const int8_t*)planes_a, ea, (const int8_t*)planes_b,
      eb, cutoff, m, n, k, na, nb, (float*)c);
  }
  else {
    return LXS_ENOVARIANT;
  }
  return (0 != result ? result : (int)hipGetLastError());
}
