/*
 * FP64 (and FP32) GEMM emulated with int8 matrix cores (Ozaki scheme I, mantissa slicing), and a
 * native FP64 matrix-core GEMM as its comparator. gfx950.
 *
 * Mathematics (the textbook scheme that samples/ozaki/kernels/ozaki1_int8.cl:432-589 and
 * ozaki_common.cl:789-825 also implement): every row i of op(A) and column j of op(B) is scaled by
 * a power of two 2^ea_i (2^eb_j) above its largest magnitude, so |x| < 1, and cut into s signed
 * 7-bit digits by truncation,
 *   x = sum_{p=0..s-1} d_p 2^(-7(p+1)) + tail,  d_p in [-127, 127],  |tail| < 2^(-7s).
 * Then (AB)(i,j) ~ 2^(ea_i + eb_j) sum_{L=0..cutoff} 2^(-7(L+2)) sum_{p+q=L} (A_p B_q)(i,j), with
 * A_p, B_q int8 digit planes. Every inner sum is an int8 GEMM accumulated exactly in integers.
 * cutoff = 2(s-1) - trim: trim 0 takes all s^2 slice pairs, trim s-1 the triangle p+q <= s-1.
 * Zero, subnormal and non-finite inputs count as zero (ozaki_common.cl:784-805).
 *
 * Kernels (new design for MI355X):
 *   oz_init     exponents := none, occupancy := 0
 *   oz_exp      per-row (A) / per-column (B) exponent: max over the row of (biased exponent - bias + 1)
 *   oz_slice    digits of 64x64 tiles into s planes [s][Rp][Kp] (k contiguous, rows padded to 256,
 *               k to 128, padding zero), so GEMM operand tiles are plain 16-byte vectors; sets one
 *               occupancy flag per plane that holds a nonzero digit
 *   oz_level_gemm  pure int8 -> int32 GEMM per work item (level L, range of its (pair, k-block) units):
 *               256x256 output tile per work-group, 4 wavefronts of 128x128 (v_mfma_i32_32x32x32_i8),
 *               operand tiles of 64 k moved into LDS by DMA, 4 stages deep; pairs whose planes are all
 *               zero are skipped. Every item writes int32 partial sums (exact: an item's units are
 *               bounded).
 *   oz_recombine   per element: S_L = exact int64 sum of the level's partials, cd = sum over
 *               L = cutoff .. 0 of 2^(-7(L+2)) S_L in FP64, C = fma(alpha, 2^(ea+eb) cd, beta C).
 *   dgemm_f64   128x128 tiles with v_mfma_f64_16x16x4f64 on FP64 operands (the comparator).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <stdlib.h>
#include <pthread.h>
#include "lxs_kernels.h"

namespace {

typedef int i4_t __attribute__((ext_vector_type(4)));
typedef int i16_t __attribute__((ext_vector_type(16)));
typedef double d4_t __attribute__((ext_vector_type(4)));

constexpr int OZ_TILE = 64; /* rows x k of a preprocessing tile */
constexpr int LT = 256;        /* int8 level GEMM: output tile (i and j) per work-group */
constexpr int LK = 128;        /* k padding of the digit planes */

/*
 * Plane layout (digits and residues alike): tiles of 256 rows x 64 k, each 16 KB contiguous (row r of a
 * tile at r * 64), tile (plane p, row block rb, k block kb) at ((p Rp / 256 + rb) Kp / 64 + kb) 16 KB.
 * One k-block of an operand tile of the int8 GEMM is then one contiguous 16 KB read.
 */
__device__ __forceinline__ long oz_plane_off(int p, int row, int k, int Rp, int Kp) {
  return (((long)p * (Rp >> 8) + (row >> 8)) * (Kp >> 6) + (k >> 6)) * 16384 + (row & 255) * 64 + (k & 63);
}
constexpr int OZ_EXP_NONE = -(1 << 30);

/* exponent ex with |v| < 2^ex for a normal finite v, OZ_EXP_NONE otherwise */
__device__ __forceinline__ int oz_expo(double v) {
  const long long bits = __double_as_longlong(v);
  const int be = (int)((bits >> 52) & 0x7FF);
  return (0 == be || 0x7FF == be) ? OZ_EXP_NONE : be - 1022;
}
__device__ __forceinline__ int oz_expo(float v) {
  const int bits = __float_as_int(v);
  const int be = (bits >> 23) & 0xFF;
  return (0 == be || 0xFF == be) ? OZ_EXP_NONE : be - 126;
}

/*
 * Preprocessing work-groups cover 64 rows x 64 k; thread t owns 16 consecutive k of one row. With rows
 * contiguous (rs == 1) a wavefront spans 64 rows (lane = row, wave = k chunk), so each of its 16 loads
 * reads 64 consecutive elements; with k contiguous (ks == 1) four lanes share a row (lane % 4 = k chunk)
 * and each lane reads 16 consecutive elements.
 */
__device__ __forceinline__ void oz_owner(long rs, int& r, int& kq) {
  const int tid = threadIdx.x;
  if (1 == rs) {
    r = tid & 63;
    kq = (tid >> 6) * 16;
  }
  else {
    r = tid >> 2;
    kq = (tid & 3) * 16;
  }
}

/* elements (row, k .. k + 15) as double; outside R x K, zero, subnormal or non-finite (in T) become 0 */
template <typename T>
__device__ __forceinline__ double oz_keep(T x) { return (OZ_EXP_NONE != oz_expo(x)) ? (double)x : 0.0; }
template <typename T>
__device__ __forceinline__ void oz_load16(double* v, const T* __restrict__ p, long rs, long ks, int row, int k, int R,
  int K) {
  const T* const q = p + (long)row * rs + (long)k * ks;
  if (row < R && k + 16 <= K) {
    if (1 == ks) { /* 16-byte loads at the element's alignment */
      constexpr int V = 16 / (int)sizeof(T);
      typedef T vec_t __attribute__((ext_vector_type(V), aligned(sizeof(T))));
#pragma unroll
      for (int i = 0; i < 16; i += V) {
        const vec_t x = *(const __attribute__((address_space(1))) vec_t*)(q + i);
#pragma unroll
        for (int j = 0; j < V; ++j) v[i + j] = oz_keep<T>(x[j]);
      }
    }
    else {
#pragma unroll
      for (int i = 0; i < 16; ++i) v[i] = oz_keep<T>(q[(long)i * ks]);
    }
  }
  else {
#pragma unroll
    for (int i = 0; i < 16; ++i) v[i] = (row < R && k + i < K) ? oz_keep<T>(q[(long)i * ks]) : 0.0;
  }
}

__global__ void __launch_bounds__(256) oz_init(int* __restrict__ e, int n, int* __restrict__ occ) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) e[i] = OZ_EXP_NONE;
  if (NULL != occ && 0 == blockIdx.x && (int)threadIdx.x < 8) occ[threadIdx.x] = 0;
}

template <typename T>
__global__ void __launch_bounds__(256) oz_exp(const T* __restrict__ p, long rs, long ks, int R, int K,
  int* __restrict__ e) {
  __shared__ int red[4][64];
  /* consecutive work-groups along the contiguous dimension */
  const int r0 = (1 == ks ? blockIdx.y : blockIdx.x) * OZ_TILE, k0 = (1 == ks ? blockIdx.x : blockIdx.y) * OZ_TILE;
  int r, kq;
  oz_owner(rs, r, kq);
  double v[16];
  oz_load16<T>(v, p, rs, ks, r0 + r, k0 + kq, R, K);
  int ex = OZ_EXP_NONE;
#pragma unroll
  for (int x = 0; x < 16; ++x) ex = max(ex, oz_expo(v[x]));
  if (1 == rs) { /* the row's four chunks are in four wavefronts */
    red[kq >> 4][r] = ex;
    __syncthreads();
    if (threadIdx.x < 64) {
      ex = max(max(red[0][r], red[1][r]), max(red[2][r], red[3][r]));
      if (OZ_EXP_NONE != ex && r0 + r < R) atomicMax(e + r0 + r, ex);
    }
  }
  else { /* four consecutive lanes */
    ex = max(ex, __shfl_xor(ex, 1));
    ex = max(ex, __shfl_xor(ex, 2));
    if (0 == (threadIdx.x & 3) && OZ_EXP_NONE != ex && r0 + r < R) atomicMax(e + r0 + r, ex);
  }
}

/* exponent pass: 64 x 64 tiles, the first grid dimension along the contiguous one */
static dim3 oz_exp_grid(long ks, int Rp, int Kp) {
  return (1 == ks) ? dim3((unsigned int)(Kp / OZ_TILE), (unsigned int)(Rp / OZ_TILE))
                   : dim3((unsigned int)(Rp / OZ_TILE), (unsigned int)(Kp / OZ_TILE));
}

/* s int8 digits of 16 consecutive k of one row, written as one 16-byte vector per plane */
template <typename T>
__global__ void __launch_bounds__(256) oz_slice(const T* __restrict__ p, long rs, long ks, int R, int K,
  const int* __restrict__ e, int s, int Rp, int Kp, int8_t* __restrict__ planes, int* __restrict__ occ) {
  const int r0 = blockIdx.x * OZ_TILE, k0 = blockIdx.y * OZ_TILE;
  const int tid = threadIdx.x;
  int r, kq;
  oz_owner(rs, r, kq);
  const int row = r0 + r;
  const int ex = (row < R ? e[row] : OZ_EXP_NONE);
  double x[16];
  oz_load16<T>(x, p, rs, ks, row, k0 + kq, R, K);
#pragma unroll
  for (int i = 0; i < 16; ++i) x[i] = (OZ_EXP_NONE != ex ? ldexp(x[i], -ex) : 0.0); /* |x| < 1, exact */
  for (int pl = 0; pl < s; ++pl) {
    union {
      int8_t b[16];
      i4_t v;
    } d;
    int nz = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const double y = x[i] * 128.0;
      const double q = trunc(y); /* |q| <= 127 */
      d.b[i] = (int8_t)q;
      nz |= (int)(0.0 != q);
      x[i] = y - q; /* exact */
    }
    *(i4_t*)(planes + oz_plane_off(pl, row, k0 + kq, Rp, Kp)) = d.v;
    if (0 != __ballot(nz) && 0 == (tid & 63) && 0 == __hip_atomic_load(occ + pl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
      atomicOr(occ + pl, 1);
    }
  }
}

/* element-wise kernels stride over the M x N elements with a 1-D grid */
constexpr int EW_BLOCKS = 1024;

template <typename T>
__global__ void __launch_bounds__(256) oz_scale_c(T* __restrict__ c, int ldc, int M, int N, double beta) {
  for (long x = blockIdx.x * 256L + threadIdx.x; x < (long)M * N; x += (long)gridDim.x * 256) {
    T* const pc = c + (x % M) + (x / M) * (long)ldc;
    *pc = (0 != beta ? (T)(beta * (double)*pc) : T(0));
  }
}

/* ---- int8 level GEMM ------------------------------------------------------------------------- */

constexpr int GT = 128;        /* output tile of the FP64 comparator */

/*
 * One work item of the int8 GEMM, in units of 64 k:
 * - Scheme 1: level L, pairs p = plo .. phi with q = L - p; unit u is pair plo + u / nkb, k-block u % nkb.
 *   The item writes its own int32 partial sums, exact because it holds at most 2080 units:
 *   2080 * 64 * 127^2 < 2^31.
 * - Scheme 2 (CRT): prime P, A and B residue planes P; unit u is k-block u. The item reduces its sums
 *   mod m_P and writes the residue as one byte (|residue| <= 128, at most 2047 units: 2047 * 64 * 2^14 < 2^31).
 */
struct oz_item_t {
  int level, u0, u1, pad;
};
constexpr int OZ_ITEMS = 128; /* items per launch, passed by value */
struct oz_items_t {
  oz_item_t it[OZ_ITEMS];
};

/* CRT moduli in use order (pairwise coprime, at most 256: symmetric residues fit int8) */
__constant__ int oz_moduli[LXS_OZ_NPRIMES_MAX] = {LXS_OZ_MODULI};

/*
 * C32[z](j, i) = sum over the item's units of sum_k A_p(i, k) B_q(j, k), stored at z * Mp * Np + j * Mp + i
 * (Scheme 2: that sum mod m_P, one byte).
 * v_mfma_i32_32x32x32_i8 with the B planes as the first operand (rows j) and the A planes as the
 * second (columns i): lane l of D holds column i = l % 32 and rows 8(r/4) + 4(l/32) + r%4, so for a
 * fixed register the 32 lanes of a half-wave store contiguous elements along i.
 * 4 wavefronts in 2 x 2, each 128 x 128 (4 x 4 MFMA tiles, 256 accumulator registers).
 * Operand tiles of 64 k (one unit) move from global memory straight into LDS by DMA
 * (global_load_lds_dwordx4), NST stages of 32 KB.
 * LDS row layout: 64 bytes per row (4 chunks of 16), chunk c of row r stored at position
 * c ^ ((r >> 2) & 3); the 16 lanes of a quarter-wave read rows r..r+15 of one chunk from 16 distinct
 * 4-bank groups. One DMA instruction fills 16 rows (1 KB): lane l writes position l % 4 of row l / 4,
 * so it loads chunk (l % 4) ^ ((row >> 2) & 3) of that row.
 * Pipeline: at stage g a wave waits until its own DMA of stage g + 1 has landed and its fragment
 * reads of stage g are complete, then meets the others at a barrier (now every wave's DMA of g + 1 has
 * landed and every wave holds stage g in registers), refills the buffer of stage g with stage g + NST,
 * reads the fragments of stage g + 1 and issues the 32 MFMAs of stage g, so the LDS reads of the next
 * stage overlap the MFMAs of this one.
 */
constexpr int QK = 64;                /* k per stage */
constexpr int QTILE = LT * QK;        /* bytes of one operand tile per stage (16 KB) */

/* operand tile of zeros: the source of the padding stages (their MFMAs add nothing) */
__device__ __attribute__((aligned(1024))) int8_t oz_zero_tile[QTILE];

__device__ __forceinline__ void oz_dma16(const int8_t* src, int8_t* lds_dst) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
    (__attribute__((address_space(3))) void*)lds_dst, 16, 0, 0);
}

/* s_waitcnt immediate (gfx9 encoding): vmcnt(vm), lgkmcnt(0), expcnt not waited for */
constexpr int oz_wait_vm_lgkm0(int vm) { return (vm & 15) | ((vm >> 4) << 14) | 0x70; }

template <int CRT, int NST>
__global__ void __launch_bounds__(256) oz_level_gemm(const int8_t* __restrict__ pa, const int8_t* __restrict__ pb,
  const int* __restrict__ occa, const int* __restrict__ occb, oz_items_t items, int s, int Mp, int Np, int Kp,
  const int8_t* __restrict__ zero, void* __restrict__ out) {
  /* one array per stage ([A|B][256 rows x 64 bytes]): distinct objects, so the compiler can tell the DMA
     targets of the stages in flight from the stage being read */
  __shared__ __attribute__((aligned(1024))) int8_t ls0[2 * QTILE], ls1[2 * QTILE], ls2[2 * QTILE], ls3[2 * QTILE];
  __shared__ __attribute__((aligned(1024))) int8_t ls4[4 == NST ? 1 : 2 * QTILE]; /* the fifth stage (NST = 5) */
  static_assert(4 == NST || 5 == NST, "the stage loop below is written for four or five buffers");
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wj = wave >> 1, wi = wave & 1;
  const int i0 = blockIdx.x * LT, j0 = blockIdx.y * LT;
  const oz_item_t it = items.it[blockIdx.z];
  const int L = it.level, nkb = Kp / QK;
  const int plo = (0 != CRT ? L : (L - s + 1 > 0 ? L - s + 1 : 0));
  const int pend = (0 != CRT ? L + 1 : (L < s - 1 ? L : s - 1) + 1);
  unsigned int oa = ~0u, ob = ~0u;
  if (0 == CRT) {
    oa = ob = 0;
    for (int p = 0; p < s; ++p) {
      oa |= (0 != occa[p] ? 1u : 0u) << p;
      ob |= (0 != occb[p] ? 1u : 0u) << p;
    }
    oa = __builtin_amdgcn_readfirstlane(oa);
    ob = __builtin_amdgcn_readfirstlane(ob);
  }
  /* pair p of the item has work: both planes hold a nonzero digit (always, for residues) */
  auto active = [&](int p) -> bool { return 0 != CRT || 0 != ((oa >> p) & (ob >> (L - p)) & 1u); };
  i16_t acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x) {
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = i16_t{0};
  }
  int nact = 0; /* active units of this item */
  for (int pp = it.u0 / nkb; pp * nkb < it.u1; ++pp) {
    const int lo = (it.u0 > pp * nkb ? it.u0 : pp * nkb), hi = (it.u1 < (pp + 1) * nkb ? it.u1 : (pp + 1) * nkb);
    if (active(plo + pp)) nact += hi - lo;
  }
  if (0 < nact) {
    /* next unit to load: plane fp, k-block fkb (wave-uniform), walked in order over the active pairs;
       ta and tb point at its operand tiles */
    int fp = plo + it.u0 / nkb, fkb = it.u0 - (it.u0 / nkb) * nkb;
    const int8_t *ta = pa, *tb = pb;
    auto locate = [&]() {
      while (fp < pend && !active(fp)) {
        ++fp;
        fkb = 0;
      }
      const int pla = (fp < pend ? fp : plo), plb = (0 != CRT ? pla : L - pla); /* tiles of 256 x 64, 16 KB each */
      ta = pa + (((long)pla * (Mp / LT) + i0 / LT) * nkb + fkb) * (long)QTILE;
      tb = pb + (((long)plb * (Np / LT) + j0 / LT) * nkb + fkb) * (long)QTILE;
    };
    locate();
    /* this lane's DMA rows 16 (wave + 4 r) + lane / 4, chunk (lane % 4) ^ ((row >> 2) & 3) (the same for all r) */
    const int drow = 16 * wave + (lane >> 2), dchunk = (lane & 3) ^ ((lane >> 4) & 3);
    const unsigned int off0 = (unsigned int)drow * QK + 16u * dchunk; /* rows 64 bytes apart in a tile */
    int g = 0;
    /* stage g into buffer buf: the next active unit while g < nact, the zero tile after that (no branch) */
    auto dma = [&](int8_t* buf) {
      const bool real = (g < nact);
      const int8_t* const ga = (real ? ta : zero);
      const int8_t* const gb = (real ? tb : zero);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        oz_dma16(ga + (off0 + r * 64 * QK), buf + (wave + 4 * r) * 1024);
        oz_dma16(gb + (off0 + r * 64 * QK), buf + QTILE + (wave + 4 * r) * 1024);
      }
    };
    /* the walk moves past the unit just loaded */
    auto advance = [&]() {
      if (g < nact) {
        if (++fkb < nkb) {
          ta += QTILE;
          tb += QTILE;
        }
        else {
          fkb = 0;
          ++fp;
          locate();
        }
      }
      ++g;
    };
    const int frow = lane & 31, fswz = (frow >> 2) & 3; /* (row >> 2) & 3 is the same for rows 32 x + frow */
    const int fpos0 = (lane >> 5) ^ fswz, fpos1 = (2 + (lane >> 5)) ^ fswz;
    const int la_off = (128 * wi + frow) * QK, lb_off = QTILE + (128 * wj + frow) * QK;
    i4_t fa[2][8], fb[2][8]; /* fragments of two stages: [k half 0 | k half 1] x 4 tiles */
    auto frags = [&](const int8_t* buf, i4_t* va, i4_t* vb) {
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        vb[x] = *(const i4_t*)(buf + lb_off + 32 * x * QK + 16 * fpos0);
        va[x] = *(const i4_t*)(buf + la_off + 32 * x * QK + 16 * fpos0);
        vb[4 + x] = *(const i4_t*)(buf + lb_off + 32 * x * QK + 16 * fpos1);
        va[4 + x] = *(const i4_t*)(buf + la_off + 32 * x * QK + 16 * fpos1);
      }
    };
    auto mfmas = [&](const i4_t* va, const i4_t* vb) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int x = 0; x < 4; ++x) {
#pragma unroll
          for (int y = 0; y < 4; ++y) {
            acc[x][y] = __builtin_amdgcn_mfma_i32_32x32x32_i8(vb[4 * h + x], va[4 * h + y], acc[x][y], 0, 0, 0);
          }
        }
      }
    };
    /* wait for own DMA of the next stage and own reads of this one, barrier */
#define OZ_SYNC() \
  do { \
    __builtin_amdgcn_s_waitcnt(oz_wait_vm_lgkm0(8 * (NST - 2))); \
    __asm__ __volatile__("" ::: "memory"); \
    __builtin_amdgcn_s_barrier(); /* no fence: it would wait for the stages still in flight too */ \
    __asm__ __volatile__("" ::: "memory"); \
  } while (0)
    /* stage g: CUR holds it (its fragments are in FC), NXT holds stage g + 1. Issue order: the MFMAs of
       stage g (their operands are in registers) interleaved with the DMA of stage g + 4 and then with the
       fragment reads of stage g + 1 */
#define OZ_STAGE(CUR, NXT, FC, FN) \
  do { \
    OZ_SYNC(); \
    dma(CUR); \
    frags(NXT, fa[FN], fb[FN]); \
    mfmas(fa[FC], fb[FC]); \
    _Pragma("unroll") for (int i_ = 0; i_ < 8; ++i_) { \
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0); /* MFMA */ \
      __builtin_amdgcn_sched_group_barrier(0x020, 1, 0); /* VMEM read (DMA) */ \
    } \
    _Pragma("unroll") for (int i_ = 0; i_ < 16; ++i_) { \
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0); /* MFMA */ \
      __builtin_amdgcn_sched_group_barrier(0x100, 1, 0); /* DS read */ \
    } \
    __builtin_amdgcn_sched_group_barrier(0x008, 8, 0); \
    advance(); \
  } while (0)
    dma(ls0);
    advance();
    dma(ls1);
    advance();
    dma(ls2);
    advance();
    if constexpr (4 == NST) {
      OZ_SYNC();
      dma(ls3);
      advance();
      frags(ls0, fa[0], fb[0]);
      for (int t = 0; t < nact; t += 4) { /* one exit: the accumulators stay in their registers */
        OZ_STAGE(ls0, ls1, 0, 1);
        OZ_STAGE(ls1, ls2, 1, 0);
        OZ_STAGE(ls2, ls3, 0, 1);
        OZ_STAGE(ls3, ls0, 1, 0);
      }
    }
    else { /* five buffers and two fragment sets: ten stages per trip */
      dma(ls3);
      advance();
      OZ_SYNC();
      dma(ls4);
      advance();
      frags(ls0, fa[0], fb[0]);
      for (int t = 0; t < nact; t += 10) {
        OZ_STAGE(ls0, ls1, 0, 1);
        OZ_STAGE(ls1, ls2, 1, 0);
        OZ_STAGE(ls2, ls3, 0, 1);
        OZ_STAGE(ls3, ls4, 1, 0);
        OZ_STAGE(ls4, ls0, 0, 1);
        OZ_STAGE(ls0, ls1, 1, 0);
        OZ_STAGE(ls1, ls2, 0, 1);
        OZ_STAGE(ls2, ls3, 1, 0);
        OZ_STAGE(ls3, ls4, 0, 1);
        OZ_STAGE(ls4, ls0, 1, 0);
      }
    }
#undef OZ_STAGE
#undef OZ_SYNC
    __builtin_amdgcn_s_waitcnt(0); /* the trailing zero-tile loads land before the work-group ends */
  }
  if (0 == CRT) {
    int* const o32 = (int*)out + (long)blockIdx.z * Mp * Np;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        const int i = i0 + 128 * wi + 32 * y + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int j = j0 + 128 * wj + 32 * x + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          o32[(long)j * Mp + i] = acc[x][y][r];
        }
      }
    }
  }
  else {
    /* v mod m for |v| < 2^31: floor(v / m) from FP64 (exact but at multiples of m, which leave r = m) */
    const int m = oz_moduli[L];
    const double rm = 1.0 / m;
    uint8_t* const o8 = (uint8_t*)out + (long)blockIdx.z * Mp * Np;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        const int i = i0 + 128 * wi + 32 * y + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int j = j0 + 128 * wj + 32 * x + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
          const int v = acc[x][y][r], q = (int)floor((double)v * rm), rr = v - q * m;
          o8[(long)j * Mp + i] = (uint8_t)(rr >= m ? rr - m : rr);
        }
      }
    }
  }
}

/* first item of every level (levels 0 .. 15), zb[L + 1] - zb[L] items per level */
struct oz_levels_t {
  int zb[17];
};

/*
 * C(i, j) = fma(alpha, 2^(ea_i + eb_j) cd, beta C(i, j)) with cd = sum over levels L = cutoff .. 0 of
 * 2^(-7(L+2)) S_L, S_L the exact (int64) sum of the level's int32 partials.
 */
template <typename T>
__global__ void __launch_bounds__(256) oz_recombine(const int* __restrict__ c32, oz_levels_t lv, int cutoff, int Mp,
  int Np, const int* __restrict__ ea, const int* __restrict__ eb, int M, int N, double alpha, double beta,
  T* __restrict__ c, int ldc) {
  for (long x = blockIdx.x * 256L + threadIdx.x; x < (long)M * N; x += (long)gridDim.x * 256) {
    const int i = (int)(x % M), j = (int)(x / M);
    const long o = (long)j * Mp + i;
    double cd = 0;
    for (int L = cutoff; L >= 0; --L) {
      long long sum = 0;
      for (int z = lv.zb[L]; z < lv.zb[L + 1]; ++z) sum += c32[(long)z * Mp * Np + o];
      cd += ldexp(1.0, -7 * (L + 2)) * (double)sum;
    }
    {
      const int ex = ea[i], ey = eb[j];
      const double v = (OZ_EXP_NONE != ex && OZ_EXP_NONE != ey) ? ldexp(cd, ex + ey) : 0.0;
      T* const pc = c + i + (long)j * ldc;
      *pc = (T)(0 != beta ? fma(alpha, v, beta * (double)*pc) : alpha * v);
    }
  }
}

/* ---- Scheme 2 (CRT) -------------------------------------------------------------------------- */

/*
 * Per modulus m (at most 256): the Barrett factor ceil(2^32 / m), exact for x < 2^18
 * (floor(x / m) = mulhi(x, rcp): the excess x (rcp - 2^32 / m) / 2^32 < 2^-14 stays below 1 / m),
 * 2^18 mod m and 2^36 mod m for the residues of 53-bit integers, and Garner's inverses
 * inv[j][i] = m_j^-1 mod m_i (j < i).
 */
struct oz_crt_tab_t {
  unsigned int m[LXS_OZ_NPRIMES_MAX], rcp[LXS_OZ_NPRIMES_MAX], k18[LXS_OZ_NPRIMES_MAX], k36[LXS_OZ_NPRIMES_MAX];
  unsigned int rcp27[LXS_OZ_NPRIMES_MAX]; /* ceil(2^39 / m) < 2^32 (m > 128): exact floor(x / m) for x < 2^27 */
  float rmf[LXS_OZ_NPRIMES_MAX];          /* 1 / m rounded to FP32 */
  unsigned int inv[LXS_OZ_NPRIMES_MAX][LXS_OZ_NPRIMES_MAX];
};
constexpr int oz_inv_mod(int a, int m) { /* a^-1 mod m (gcd(a, m) = 1), extended Euclid */
  int t = 0, nt = 1, r = m, nr = a % m;
  while (0 != nr) {
    const int q = r / nr, t2 = t - q * nt, r2 = r - q * nr;
    t = nt, nt = t2, r = nr, nr = r2;
  }
  return (0 > t ? t + m : t);
}
constexpr oz_crt_tab_t oz_crt_make() {
  constexpr int mod[LXS_OZ_NPRIMES_MAX] = {LXS_OZ_MODULI};
  oz_crt_tab_t t{};
  for (int i = 0; i < LXS_OZ_NPRIMES_MAX; ++i) {
    const unsigned long long m = (unsigned long long)mod[i];
    t.m[i] = (unsigned int)m;
    t.rcp[i] = (unsigned int)(((1ull << 32) + m - 1) / m);
    t.rcp27[i] = (unsigned int)(((1ull << 39) + m - 1) / m);
    t.rmf[i] = 1.0f / (float)m;
    t.k18[i] = (unsigned int)((1ull << 18) % m);
    t.k36[i] = (unsigned int)((1ull << 36) % m);
    for (int j = 0; j < i; ++j) t.inv[j][i] = (unsigned int)oz_inv_mod(mod[j] % mod[i], mod[i]);
  }
  return t;
}
__constant__ oz_crt_tab_t oz_crt = oz_crt_make();

/*
 * Residue tables: byte i (i = 0..7) of kb[w][p] is 2^(8 i) w_p mod m_p, where w_p is 1 (w = 0: operand B) or
 * Garner's weight of the reconstruction for np = w primes, ((M / m_p) mod m_p)^-1 mod m_p (operand A: its
 * residues carry the weight, so the int8 products modulo m_p are already the CRT coefficients c_p). lo is
 * the least symmetric residue, (m + 1) / 2 - m; r31 = ceil(2^31 / m) (m > 128).
 */
struct oz_res_tab_t {
  unsigned int kb[LXS_OZ_NPRIMES_MAX + 1][LXS_OZ_NPRIMES_MAX][2];
  int lo[LXS_OZ_NPRIMES_MAX];
  unsigned int r31[LXS_OZ_NPRIMES_MAX];
};
constexpr oz_res_tab_t oz_res_make() {
  constexpr int mod[LXS_OZ_NPRIMES_MAX] = {LXS_OZ_MODULI};
  oz_res_tab_t t{};
  for (int p = 0; p < LXS_OZ_NPRIMES_MAX; ++p) {
    const unsigned long long m = (unsigned long long)mod[p];
    t.lo[p] = (mod[p] + 1) / 2 - mod[p];
    t.r31[p] = (unsigned int)(((1ull << 31) + m - 1) / m);
  }
  for (int np = 0; np <= LXS_OZ_NPRIMES_MAX; ++np) {
    unsigned __int128 mm = 1;
    for (int p = 0; p < np; ++p) mm *= (unsigned __int128)mod[p];
    for (int p = 0; p < LXS_OZ_NPRIMES_MAX; ++p) {
      const unsigned long long m = (unsigned long long)mod[p];
      const unsigned long long w = (0 < np && p < np)
                                     ? (unsigned long long)oz_inv_mod((int)((mm / m) % (unsigned __int128)m), (int)m)
                                     : 1ull;
      unsigned long long pw = 1; /* 2^(8 i) mod m */
      for (int i = 0; i < 8; ++i) {
        t.kb[np][p][i >> 2] |= (unsigned int)((pw * w) % m) << (8 * (i & 3));
        pw = (pw * 256) % m;
      }
    }
  }
  return t;
}
__constant__ oz_res_tab_t oz_res = oz_res_make();

__device__ __forceinline__ unsigned int oz_mod18(unsigned int x, int p) { /* x mod m_p for x < 2^18 */
  return x - __umul24(__umulhi(x, oz_crt.rcp[p]), oz_crt.m[p]);
}
/* x mod m_p for x < 2^27: the excess x (rcp27 - 2^39 / m) / 2^39 < 2^-12 stays below 1 / m */
__device__ __forceinline__ unsigned int oz_mod27(unsigned int x, int p) {
  return x - __umul24(__umulhi(x, oz_crt.rcp27[p]) >> 7, oz_crt.m[p]);
}

/*
 * Residue planes of one 64 x 64 tile: every element v of row r with the row's exponent E (|v| < 2^E)
 * becomes the integer a = trunc(|v| 2^(mb - t - E)) (mb = 53 for FP64, 24 for FP32; t trimmed bits);
 * plane p holds the symmetric residue of sign(v) a w_p mod m_p, in [lo_p, lo_p + m_p), as int8 (w_p: 1, or
 * the reconstruction weight for weight = np > 0, see oz_res_tab_t). A thread owns 16 consecutive k of one
 * row and writes one 16-byte vector per plane.
 * Per plane and element: x = sum of the 7 bytes of a times 2^(8 i) w mod m (two byte dot products,
 * x < 7 * 255^2 < 2^19, x = a w mod m), y = x - lo or m 2^12 - x - lo by the sign (so y = sign a w - lo
 * mod m and 0 <= y < 2^21; the sign enters through an all-ones mask, -x = (x ^ ~0) + 1), y mod m by one 24-bit Barrett step (floor(y / m) =
 * (2 y ceil(2^31 / m)) >> 32: the excess below 2^-10 cannot reach the next integer, y / m being a multiple
 * of 1 / m >= 2^-8), and the residue is lo + (y mod m).
 */
template <typename T>
__global__ void __launch_bounds__(256) oz_residues(const T* __restrict__ p, long rs, long ks, int R, int K,
  const int* __restrict__ e, int np, int weight, int mb, int trim, int Rp, int Kp, int8_t* __restrict__ planes) {
  const int r0 = blockIdx.x * OZ_TILE, k0 = blockIdx.y * OZ_TILE;
  /* four lanes per row (k chunks of 16), 16 rows per wavefront, for either operand layout: each lane's 16
     residue bytes of a plane continue its neighbour's, so a wavefront stores 1 KB contiguous per plane */
  const int r = (int)threadIdx.x >> 2, kq = ((int)threadIdx.x & 3) * 16;
  const int row = r0 + r;
  const int ex = (row < R ? e[row] : OZ_EXP_NONE);
  double t[16];
  oz_load16<T>(t, p, rs, ks, row, k0 + kq, R, K);
  unsigned int alo[16], ahi[16]; /* a = ahi 2^32 + alo, a < 2^53 */
  unsigned int nm[16];           /* all ones for negative v, else 0 */
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const double v = t[i];
    const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    const int ev = oz_expo(v); /* |v| < 2^ev */
    const unsigned long long mant = (bits & 0xFFFFFFFFFFFFFull) | (1ull << 52);
    const int shift = (53 - mb) + trim + (ex - ev); /* a = mant >> shift */
    const unsigned long long a = (OZ_EXP_NONE != ev && OZ_EXP_NONE != ex && shift < 53) ? (mant >> shift) : 0ull;
    alo[i] = (unsigned int)a;
    ahi[i] = (unsigned int)(a >> 32);
    nm[i] = 0u - (unsigned int)(bits >> 63);
  }
  const int w = (0 != weight ? np : 0);
#pragma unroll
  for (int pl = 0; pl < LXS_OZ_NPRIMES_MAX; ++pl) {
    if (pl < np) {
      const unsigned int m = oz_crt.m[pl], klo = oz_res.kb[w][pl][0], khi = oz_res.kb[w][pl][1], r31 = oz_res.r31[pl];
      const int lo = oz_res.lo[pl];
      const unsigned int cbig = (m << 12) + 1u, clo = (unsigned int)-lo; /* -x = (x ^ ~0) + 1 */
      union {
        int8_t b[16];
        i4_t v;
      } d;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const unsigned int x = __builtin_amdgcn_udot4(alo[i], klo, __builtin_amdgcn_udot4(ahi[i], khi, 0u, false), false);
        const unsigned int y = (x ^ nm[i]) + (nm[i] & cbig) + clo; /* x - lo, or m 2^12 - x - lo */
        /* both factors below 2^24 (masks state it for the compiler: v_mul_hi_u32_u24) */
        const unsigned int q = (unsigned int)(((unsigned long long)((2u * y) & 0xFFFFFFu) * (r31 & 0xFFFFFFu)) >> 32);
        /* lo + y mod m = lo + y - q m, of which only the low byte is kept: lo + y + q (256 - m) */
        d.b[i] = (int8_t)(__umul24(q, 256u - m) + y + (unsigned int)lo);
      }
      *(i4_t*)(planes + oz_plane_off(pl, row, k0 + kq, Rp, Kp)) = d.v;
    }
  }
}

/*
 * Constants of the reconstruction for every prime count np (M = product of the first np moduli),
 * computed at compile time: w = ((M / m_p) mod m_p)^-1 mod m_p, M / m_p in 16-bit limbs (least
 * significant first), M in 16-bit limbs, M / 2 and M as {high, low} 64-bit words.
 */
struct oz_crt_rec_t {
  unsigned int w[LXS_OZ_NPRIMES_MAX + 1][LXS_OZ_NPRIMES_MAX];
  unsigned int l16[LXS_OZ_NPRIMES_MAX + 1][LXS_OZ_NPRIMES_MAX][8];
  unsigned int l16p[LXS_OZ_NPRIMES_MAX + 1][LXS_OZ_NPRIMES_MAX / 2][8]; /* limbs of primes 2h, 2h + 1 packed */
  unsigned int m16[LXS_OZ_NPRIMES_MAX + 1][8];
  unsigned long long hm[LXS_OZ_NPRIMES_MAX + 1][2], mm[LXS_OZ_NPRIMES_MAX + 1][2];
  double rm[LXS_OZ_NPRIMES_MAX]; /* 1 / m_p */
};
constexpr oz_crt_rec_t oz_crt_rec_make() {
  constexpr int mod[LXS_OZ_NPRIMES_MAX] = {LXS_OZ_MODULI};
  oz_crt_rec_t t{};
  for (int p = 0; p < LXS_OZ_NPRIMES_MAX; ++p) t.rm[p] = 1.0 / mod[p];
  for (int np = 1; np <= LXS_OZ_NPRIMES_MAX; ++np) {
    unsigned __int128 m = 1;
    for (int p = 0; p < np; ++p) m *= (unsigned __int128)mod[p];
    for (int k = 0; k < 8; ++k) t.m16[np][k] = (unsigned int)((m >> (16 * k)) & 0xFFFF);
    t.mm[np][0] = (unsigned long long)(m >> 64);
    t.mm[np][1] = (unsigned long long)m;
    t.hm[np][0] = (unsigned long long)((m >> 1) >> 64);
    t.hm[np][1] = (unsigned long long)(m >> 1);
    for (int p = 0; p < np; ++p) {
      const unsigned __int128 mp = m / (unsigned __int128)mod[p];
      t.w[np][p] = (unsigned int)oz_inv_mod((int)(mp % (unsigned __int128)mod[p]), mod[p]);
      for (int k = 0; k < 8; ++k) t.l16[np][p][k] = (unsigned int)((mp >> (16 * k)) & 0xFFFF);
    }
    for (int h = 0; h < LXS_OZ_NPRIMES_MAX / 2; ++h) {
      for (int k = 0; k < 8; ++k) t.l16p[np][h][k] = t.l16[np][2 * h][k] | (t.l16[np][2 * h + 1][k] << 16);
    }
  }
  return t;
}
__constant__ oz_crt_rec_t oz_rec = oz_crt_rec_make();

typedef unsigned short oz_u16x2_t __attribute__((ext_vector_type(2)));

/* 128-bit two's complement value (hi signed, lo unsigned) */
struct oz_i128_t {
  long long hi;
  unsigned long long lo;
};
__device__ __forceinline__ oz_i128_t oz_add(oz_i128_t a, long long bhi, unsigned long long blo) {
  oz_i128_t r;
  r.lo = a.lo + blo;
  r.hi = a.hi + bhi + (r.lo < blo ? 1 : 0);
  return r;
}
__device__ __forceinline__ oz_i128_t oz_sub(oz_i128_t a, long long bhi, unsigned long long blo) {
  oz_i128_t r;
  r.lo = a.lo - blo;
  r.hi = a.hi - bhi - (a.lo < blo ? 1 : 0);
  return r;
}

/*
 * C(i, j) = fma(alpha, 2^(ea_i + eb_j + sh0) x, beta C(i, j)) where x is the integer of the residues
 * r_p = sum of the prime's residue bytes mod m_p, |x| < M / 2. Direct CRT: with c_p = r_p w_p mod m_p,
 * x_raw = sum_p c_p (M / m_p) = x (mod M) and x_raw / M = sum_p c_p / m_p, so x = x_raw - q M for
 * q = round(sum_p c_p / m_p) (FP32: |x_raw - q M| < (1/2 + 2^-14) M), then at most one exact step of M
 * into [-M/2, M/2). x_raw is summed in eight 16-bit limbs of 32-bit accumulators, two primes per
 * v_dot2_u32_u16 (c_p (M / m_p)_k < 2^24, 16 terms < 2^28: no carries until the end). |x| is rounded to FP64 once (round to nearest even). Four consecutive
 * elements of a column per thread.
 */
template <typename T>
__global__ void __launch_bounds__(256) oz_crt_recombine(const uint8_t* __restrict__ r8, oz_levels_t lv, int np,
  int Mp, int Np, const int* __restrict__ ea, const int* __restrict__ eb, int M, int N, int sh0, double alpha,
  double beta, T* __restrict__ c, int ldc) {
  /* one thread per 4 elements, no loop: the constants are read where they are used (scalar operands) */
  const unsigned int mq = (unsigned int)Mp >> 2, total = mq * (unsigned int)N; /* < 2^32 (launcher) */
  const long zs = (long)Mp * Np;
  const long long hm_hi = (long long)oz_rec.hm[np][0], m_hi = (long long)oz_rec.mm[np][0];
  const unsigned long long hm_lo = oz_rec.hm[np][1], m_lo = oz_rec.mm[np][1];
  {
    const unsigned int x = blockIdx.x * 256u + threadIdx.x;
    if (x >= total) return;
    const unsigned int jq = x / mq;
    const int i0 = (int)(4 * (x - jq * mq)), j = (int)jq;
    if (i0 >= M) return;
    const unsigned int lo = (unsigned int)j * (unsigned int)Mp + (unsigned int)i0; /* < Mp Np < 2^32 */
    unsigned int acc[4][8];
    float f[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      f[e] = 0.0f;
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[e][k] = 0;
    }
    /* the first item's bytes of every prime, loaded before any is used (primes past np reload the last) */
    unsigned int wv[LXS_OZ_NPRIMES_MAX];
#pragma unroll
    for (int p = 0; p < LXS_OZ_NPRIMES_MAX; ++p) {
      wv[p] = *(const unsigned int*)(r8 + lv.zb[p < np ? p : np - 1] * zs + lo); /* uniform base + 32-bit offset */
    }
#pragma unroll
    for (int p2 = 0; p2 < LXS_OZ_NPRIMES_MAX; p2 += 2) {
      if (p2 < np) {
        unsigned int cpair[4] = {0, 0, 0, 0}; /* c_p of primes p2 (low half) and p2 + 1 (high half) */
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int p = p2 + h;
          if (p < np) {
            const unsigned int mp = oz_crt.m[p];
            const float rmp = oz_crt.rmf[p];
            /* r_p = sum of the items' bytes (each < m_p) mod m_p: min(t, t - m) is t mod m for t < 2 m */
            int z = lv.zb[p];
            unsigned int w = wv[p], sv[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) sv[e] = (w >> (8 * e)) & 0xFF;
            for (++z; z < lv.zb[p + 1]; ++z) {
              w = *(const unsigned int*)(r8 + z * zs + lo);
#pragma unroll
              for (int e = 0; e < 4; ++e) {
                const unsigned int t = sv[e] + ((w >> (8 * e)) & 0xFF);
                sv[e] = min(t, t - mp);
              }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              /* the A residues carry w_p (oz_res_tab_t), so r_p is already c_p = r w_p mod m_p */
              const unsigned int cp = sv[e];
              f[e] = fmaf((float)cp, rmp, f[e]); /* sum_p c_p / m_p < np: absolute error below 2^-14 */
              cpair[e] |= cp << (16 * h);
            }
          }
        }
        /* limb k += c_p (M / m_p)_k + c_p+1 (M / m_p+1)_k: products < 2^24, 16 terms < 2^28 */
#pragma unroll
        for (int e = 0; e < 4; ++e) {
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            acc[e][k] = __builtin_amdgcn_udot2(__builtin_bit_cast(oz_u16x2_t, cpair[e]),
              __builtin_bit_cast(oz_u16x2_t, oz_rec.l16p[np][p2 >> 1][k]), acc[e][k], false);
          }
        }
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int i = i0 + e;
      if (i < M) {
        const int q = (int)floorf(f[e] + 0.5f); /* 0 .. np; off by one at most, which the step below absorbs */
        /* x_raw - q M in 16-bit limbs, carries as arithmetic shifts; the top limb keeps the sign */
        unsigned long long lo = 0, hi = 0;
        int carry = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
          const int t = (int)acc[e][k] - q * (int)oz_rec.m16[np][k] + carry;
          carry = t >> 16;
          if (k < 4) lo |= (unsigned long long)(t & 0xFFFF) << (16 * k);
          else hi |= (unsigned long long)(t & 0xFFFF) << (16 * (k - 4));
        }
        oz_i128_t v;
        v.hi = (long long)hi; /* |x_raw - q M| < M < 2^126: bits 126, 127 are the sign */
        v.lo = lo;
        { /* into [-M/2, M/2) */
          const oz_i128_t up = oz_sub(v, hm_hi, hm_lo); /* v - M/2 >= 0: v -= M */
          const oz_i128_t dn = oz_add(v, hm_hi, hm_lo); /* v + M/2 < 0: v += M */
          if (0 <= up.hi) v = oz_sub(v, m_hi, m_lo);
          else if (0 > dn.hi) v = oz_add(v, m_hi, m_lo);
        }
        const bool neg = (0 > v.hi);
        unsigned long long uh = (unsigned long long)v.hi, ul = v.lo;
        if (neg) { /* |x| */
          ul = ~ul + 1;
          uh = ~uh + (0 == ul ? 1 : 0);
        }
        double d; /* |x| rounded to FP64: the top 64 bits with a sticky bit below, converted with one rounding */
        {
          unsigned long long top = ul;
          int scale = 0;
          if (0 != uh) {
            const int lz = __clzll((long long)uh);
            top = (0 != lz ? (uh << lz) | (ul >> (64 - lz)) : uh) | ((0 != (ul << lz)) ? 1ull : 0ull);
            scale = 64 - lz;
          }
          d = ldexp(fma((double)(unsigned int)(top >> 32), 4294967296.0, (double)(unsigned int)top), scale);
        }
        const int ex = ea[i], ey = eb[j];
        const double val = (OZ_EXP_NONE != ex && OZ_EXP_NONE != ey) ? ldexp(neg ? -d : d, ex + ey + sh0) : 0.0;
        T* const pc = c + i + (long)j * ldc;
        *pc = (T)(0 != beta ? fma(alpha, val, beta * (double)*pc) : alpha * val);
      }
    }
  }
}

/* ---- native FP64 GEMM (comparator) ----------------------------------------------------------- */

constexpr int DK = 16; /* k per LDS stage */

/* v_mfma_f64_16x16x4f64: A lane l -> A[l%16][l/16], B lane l -> B[l/16][l%16], D lane l reg r -> (l/16 + 4r, l%16) */
__global__ void __launch_bounds__(256) dgemm_f64(const double* __restrict__ a, long ars, long aks,
  const double* __restrict__ b, long bks, long bcs, int M, int N, int K, double alpha, double beta,
  double* __restrict__ c, int ldc) {
  __shared__ double la[2][DK][GT], lb[2][DK][GT]; /* [buffer][k][row / column] */
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wr = wave >> 1, wc = wave & 1;
  const int row0 = blockIdx.x * GT, col0 = blockIdx.y * GT;
  const int nkb = (K + DK - 1) / DK;
  d4_t acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x) {
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = d4_t{0, 0, 0, 0};
  }
  /* staging: 128 x 16 elements per operand, 8 per thread; element e = tid + 256 x */
  double ra[8], rb[8];
  auto fetch = [&](int kb) {
    const int k0 = kb * DK;
#pragma unroll
    for (int x = 0; x < 8; ++x) {
      const int e = tid + 256 * x;
      const int r = (1 == ars ? (e & 127) : (e >> 4)), k = (1 == ars ? (e >> 7) : (e & 15));
      const int gi = row0 + r, gk = k0 + k;
      ra[x] = (gi < M && gk < K) ? a[(long)gi * ars + (long)gk * aks] : 0.0;
      const int cc = (1 == bcs ? (e & 127) : (e >> 4)), kc = (1 == bcs ? (e >> 7) : (e & 15));
      const int gj = col0 + cc, gk2 = k0 + kc;
      rb[x] = (gj < N && gk2 < K) ? b[(long)gk2 * bks + (long)gj * bcs] : 0.0;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int x = 0; x < 8; ++x) {
      const int e = tid + 256 * x;
      if (1 == ars) la[buf][e >> 7][e & 127] = ra[x];
      else la[buf][e & 15][e >> 4] = ra[x];
      if (1 == bcs) lb[buf][e >> 7][e & 127] = rb[x];
      else lb[buf][e & 15][e >> 4] = rb[x];
    }
  };
  fetch(0);
  store(0);
  __syncthreads();
  for (int kb = 0; kb < nkb; ++kb) {
    const int buf = kb & 1;
    if (kb + 1 < nkb) fetch(kb + 1);
#pragma unroll
    for (int ks = 0; ks < DK; ks += 4) {
      double fa[4], fb[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        fa[x] = la[buf][ks + (lane >> 4)][64 * wr + 16 * x + (lane & 15)];
        fb[x] = lb[buf][ks + (lane >> 4)][64 * wc + 16 * x + (lane & 15)];
      }
#pragma unroll
      for (int x = 0; x < 4; ++x) {
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[x], fb[y], acc[x][y], 0, 0, 0);
      }
    }
    if (kb + 1 < nkb) store(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 4; ++x) {
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int j = col0 + 64 * wc + 16 * y + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = row0 + 64 * wr + 16 * x + (lane >> 4) + 4 * r;
        if (i < M && j < N) {
          double* const pc = c + i + (long)j * ldc;
          *pc = (0 != beta ? fma(alpha, acc[x][y][r], beta * *pc) : alpha * acc[x][y][r]);
        }
      }
    }
  }
}

/* ---- complex 3M helpers ---------------------------------------------------------------------- */

template <typename T>
__global__ void __launch_bounds__(256) oz_csplit(const T* __restrict__ z, int ld, int rows, int cols, int conj, T* __restrict__ re,
  T* __restrict__ im, T* __restrict__ sum) {
  for (long o = blockIdx.x * 256L + threadIdx.x; o < (long)rows * cols; o += (long)gridDim.x * 256) {
    const long zi = 2 * ((o % rows) + (o / rows) * (long)ld);
    const T x = z[zi], y = (0 != conj ? -z[zi + 1] : z[zi + 1]);
    re[o] = x;
    im[o] = y;
    sum[o] = x + y;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) oz_ccombine(const T* __restrict__ t1, const T* __restrict__ t2,
  const T* __restrict__ t3, int M, int N, double ar, double ai, double br, double bi, T* __restrict__ c, int ldc) {
  for (long o = blockIdx.x * 256L + threadIdx.x; o < (long)M * N; o += (long)gridDim.x * 256) {
    const long i = o % M, j = o / M;
    const double pr = (double)t1[o] - (double)t2[o], pi = (double)t3[o] - (double)t1[o] - (double)t2[o];
    T* const pc = c + 2 * (i + j * ldc);
    double cr = ar * pr - ai * pi, ci = ar * pi + ai * pr;
    if (0 != br || 0 != bi) {
      const double xr = (double)pc[0], xi = (double)pc[1];
      cr += br * xr - bi * xi;
      ci += br * xi + bi * xr;
    }
    pc[0] = (T)cr;
    pc[1] = (T)ci;
  }
}

} /* namespace */

/* digit planes, exponents and occupancy of one operand: element (r, k) at p[r * rs + k * ks] */
extern "C" int lxs_oz_preprocess(int dtype, const void* p, long rs, long ks, int R, int K, int s, int Rp, int Kp,
  int* e, int* occ, signed char* planes, void* hip_stream) {
  const hipStream_t st = (hipStream_t)hip_stream;
  const dim3 egrid = oz_exp_grid(ks, Rp, Kp), grid((unsigned int)(Rp / OZ_TILE), (unsigned int)(Kp / OZ_TILE));
  if (0 != Rp % LT || 0 != Kp % LK || Rp < R || Kp < K || 0 >= Kp || 1 > s || 8 < s) return LXS_ELIMIT;
  if (!((1 == rs) || (1 == ks))) return LXS_ELIMIT;
  oz_init<<<(Rp + 255) / 256, 256, 0, st>>>(e, Rp, occ);
  if (3 == dtype) {
    oz_exp<double><<<egrid, 256, 0, st>>>((const double*)p, rs, ks, R, K, e);
    oz_slice<double><<<grid, 256, 0, st>>>((const double*)p, rs, ks, R, K, e, s, Rp, Kp, planes, occ);
  }
  else if (1 == dtype) {
    oz_exp<float><<<egrid, 256, 0, st>>>((const float*)p, rs, ks, R, K, e);
    oz_slice<float><<<grid, 256, 0, st>>>((const float*)p, rs, ks, R, K, e, s, Rp, Kp, planes, occ);
  }
  else {
    return LXS_ENOVARIANT;
  }
  return (int)hipGetLastError();
}

/* device address of the zero tile on the current device (one lookup per device) */
static const int8_t* oz_zero_ptr() {
  static const int8_t* ptr[64];
  int dev = 0;
  if (hipSuccess != hipGetDevice(&dev) || 0 > dev || 64 <= dev) return NULL;
  if (NULL == ptr[dev]) {
    void* p = NULL;
    if (hipSuccess != hipGetSymbolAddress(&p, HIP_SYMBOL(oz_zero_tile))) return NULL;
    ptr[dev] = (const int8_t*)p;
  }
  return ptr[dev];
}

/*
 * Device scratch of ozaki_gemm_noalloc: one array of LXS_OZ_STATIC_BYTES per device, held until the
 * process ends; every call on a device shares it, so calls that overlap on one device race. On the host
 * the array's name only stands in for the symbol: its device address comes from hipGetSymbolAddress,
 * looked up once per device with that device current, and only that address reaches kernels.
 */
static __device__ __align__(256) unsigned char scratch[LXS_OZ_STATIC_BYTES];

extern "C" int lxs_oz_static_scratch(void** ptr) {
  static pthread_mutex_t lock = PTHREAD_MUTEX_INITIALIZER;
  static void* addr[64];
  int dev = 0, result = 0;
  if (NULL == ptr) return LXS_ELIMIT;
  *ptr = NULL;
  result = (int)hipGetDevice(&dev);
  if (0 != result) return result;
  if (0 > dev || 64 <= dev) return LXS_ELIMIT;
  pthread_mutex_lock(&lock);
  if (NULL == addr[dev]) {
    void* p = NULL;
    result = (int)hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (0 == result && NULL == p) result = LXS_ELIMIT;
    if (0 == result) addr[dev] = p;
  }
  *ptr = addr[dev];
  pthread_mutex_unlock(&lock);
  return result;
}

/* items: host array of nitems x {level, u0, u1, pad}; item z writes at out + z * Mp * Np (int32 partials,
   or residue bytes with crt); planes row-padded to a multiple of 256 and k-padded to a multiple of 128 */
extern "C" int lxs_oz_level_gemm(int crt, const signed char* pa, const signed char* pb, const int* occa,
  const int* occb, const int* items, int nitems, int s, int Mp, int Np, int Kp, void* out, void* hip_stream) {
  const int8_t* const zero = oz_zero_ptr();
  const int nkb = Kp / QK;
  if (0 != Mp % LT || 0 != Np % LT || 0 != Kp % LK || 0 >= Kp || (1 << 22) < Kp || 0 >= nitems || NULL == zero) {
    return LXS_ELIMIT;
  }
  if (0 == crt ? (1 > s || 8 < s || NULL == occa || NULL == occb) : (1 > s || LXS_OZ_NPRIMES_MAX < s)) return LXS_ELIMIT;
  for (int z = 0; z < nitems; ++z) {
    const int L = items[4 * z], u0 = items[4 * z + 1], u1 = items[4 * z + 2];
    if (0 == crt) { /* units of 64 k within the level's pairs, at most 2080 per item */
      const int plo = (L - s + 1 > 0 ? L - s + 1 : 0), phi = (L < s - 1 ? L : s - 1);
      if (0 > L || 2 * (s - 1) < L || 0 > u0 || u0 >= u1 || 2080 < u1 - u0 || (phi - plo + 1) * nkb < u1) return LXS_ELIMIT;
    }
    else if (0 > L || s <= L || 0 > u0 || u0 >= u1 || 2047 < u1 - u0 || nkb < u1) { /* k-blocks of prime L */
      return LXS_ELIMIT;
    }
  }
  /* LDS stages: five (ten-stage trips) when every item is long enough for the padding to vanish, else four;
     LIBXSTREAM_OZAKI_STAGES=4|5 overrides */
  int nst = 5;
  {
    const char* const s_env = getenv("LIBXSTREAM_OZAKI_STAGES");
    const int env = (NULL != s_env ? atoi(s_env) : 0);
    for (int z = 0; z < nitems; ++z) {
      if (40 > items[4 * z + 2] - items[4 * z + 1]) nst = 4;
    }
    if (4 == env || 5 == env) nst = env;
  }
  for (int z0 = 0; z0 < nitems; z0 += OZ_ITEMS) {
    const int nz = (nitems - z0 < OZ_ITEMS ? nitems - z0 : OZ_ITEMS);
    const dim3 grid((unsigned int)(Mp / LT), (unsigned int)(Np / LT), (unsigned int)nz);
    oz_items_t arg;
    for (int z = 0; z < OZ_ITEMS; ++z) {
      const int* const src = items + 4 * (z0 + (z < nz ? z : 0));
      arg.it[z].level = src[0];
      arg.it[z].u0 = src[1];
      arg.it[z].u1 = src[2];
      arg.it[z].pad = 0;
    }
    const hipStream_t st = (hipStream_t)hip_stream;
    void* const o = (0 == crt ? (void*)((int*)out + (long)z0 * Mp * Np) : (void*)((uint8_t*)out + (long)z0 * Mp * Np));
    const int8_t *a8 = (const int8_t*)pa, *b8 = (const int8_t*)pb;
    if (0 == crt && 4 == nst) oz_level_gemm<0, 4><<<grid, 256, 0, st>>>(a8, b8, occa, occb, arg, s, Mp, Np, Kp, zero, o);
    else if (0 == crt) oz_level_gemm<0, 5><<<grid, 256, 0, st>>>(a8, b8, occa, occb, arg, s, Mp, Np, Kp, zero, o);
    else if (4 == nst) oz_level_gemm<1, 4><<<grid, 256, 0, st>>>(a8, b8, occa, occb, arg, s, Mp, Np, Kp, zero, o);
    else oz_level_gemm<1, 5><<<grid, 256, 0, st>>>(a8, b8, occa, occb, arg, s, Mp, Np, Kp, zero, o);
  }
  return (int)hipGetLastError();
}

extern "C" int lxs_oz_crt_preprocess(int dtype, const void* p, long rs, long ks, int R, int K, int np, int weight,
  int trim, int Rp, int Kp, int* e, signed char* planes, void* hip_stream) {
  const hipStream_t st = (hipStream_t)hip_stream;
  const dim3 egrid = oz_exp_grid(ks, Rp, Kp), grid((unsigned int)(Rp / OZ_TILE), (unsigned int)(Kp / OZ_TILE));
  const int mb = (3 == dtype ? 53 : 24);
  if (0 != Rp % LT || 0 != Kp % LK || Rp < R || Kp < K || 0 >= Kp || 1 > np || LXS_OZ_NPRIMES_MAX < np) return LXS_ELIMIT;
  if (!((1 == rs) || (1 == ks)) || 0 > trim || mb <= trim) return LXS_ELIMIT;
  oz_init<<<(Rp + 255) / 256, 256, 0, st>>>(e, Rp, NULL);
  if (3 == dtype) {
    oz_exp<double><<<egrid, 256, 0, st>>>((const double*)p, rs, ks, R, K, e);
    oz_residues<double><<<grid, 256, 0, st>>>((const double*)p, rs, ks, R, K, e, np, weight, mb, trim, Rp, Kp,
      (int8_t*)planes);
  }
  else if (1 == dtype) {
    oz_exp<float><<<egrid, 256, 0, st>>>((const float*)p, rs, ks, R, K, e);
    oz_residues<float><<<grid, 256, 0, st>>>((const float*)p, rs, ks, R, K, e, np, weight, mb, trim, Rp, Kp,
      (int8_t*)planes);
  }
  else {
    return LXS_ENOVARIANT;
  }
  return (int)hipGetLastError();
}

/* zb: first item per prime, np + 1 entries; C = alpha 2^(ea + eb + sh0) x + beta C */
extern "C" int lxs_oz_crt_recombine(int dtype, const unsigned char* r8, const int* zb, int np, int Mp, int Np,
  const int* ea, const int* eb, int M, int N, int sh0, double alpha, double beta, void* c, int ldc, void* hip_stream) {
  oz_levels_t lv;
  const long blocks = ((long)(Mp / 4) * N + 255) / 256; /* one thread per 4 elements */
  const unsigned int grid = (unsigned int)blocks;
  if (1 > np || LXS_OZ_NPRIMES_MAX < np || Mp < M || Np < N || 0 != Mp % 4) return LXS_ELIMIT;
  if (4294967295.0 - 256.0 < (double)(Mp / 4) * N) return LXS_ELIMIT; /* 32-bit element index */
  if (0 >= M || 0 >= N) return 0;
  for (int p = 0; p < 17; ++p) lv.zb[p] = (p <= np ? zb[p] : zb[np]);
  if (3 == dtype) {
    oz_crt_recombine<double><<<grid, 256, 0, (hipStream_t)hip_stream>>>((const uint8_t*)r8, lv, np, Mp, Np, ea, eb, M, N,
      sh0, alpha, beta, (double*)c, ldc);
  }
  else if (1 == dtype) {
    oz_crt_recombine<float><<<grid, 256, 0, (hipStream_t)hip_stream>>>((const uint8_t*)r8, lv, np, Mp, Np, ea, eb, M, N,
      sh0, alpha, beta, (float*)c, ldc);
  }
  else {
    return LXS_ENOVARIANT;
  }
  return (int)hipGetLastError();
}

/* zb: first item per level, cutoff + 2 entries */
extern "C" int lxs_oz_recombine(int dtype, const int* c32, const int* zb, int cutoff, int Mp, int Np, const int* ea,
  const int* eb, int M, int N, double alpha, double beta, void* c, int ldc, void* hip_stream) {
  oz_levels_t lv;
  const long blocks = ((long)M * N + 255) / 256;
  const unsigned int grid = (unsigned int)(blocks < 4096 ? blocks : 4096);
  if (0 > cutoff || 14 < cutoff || Mp < M || Np < N) return LXS_ELIMIT;
  if (0 >= M || 0 >= N) return 0;
  for (int L = 0; L < 17; ++L) lv.zb[L] = (L <= cutoff + 1 ? zb[L] : zb[cutoff + 1]);
  if (3 == dtype) {
    oz_recombine<double><<<grid, 256, 0, (hipStream_t)hip_stream>>>(c32, lv, cutoff, Mp, Np, ea, eb, M, N, alpha, beta,
      (double*)c, ldc);
  }
  else if (1 == dtype) {
    oz_recombine<float><<<grid, 256, 0, (hipStream_t)hip_stream>>>(c32, lv, cutoff, Mp, Np, ea, eb, M, N, alpha, beta,
      (float*)c, ldc);
  }
  else {
    return LXS_ENOVARIANT;
  }
  return (int)hipGetLastError();
}

/* C = beta C (alpha = 0 or K = 0) */
extern "C" int lxs_oz_scale(int dtype, void* c, int ldc, int M, int N, double beta, void* hip_stream) {
  const long blocks = ((long)M * N + 255) / 256;
  const unsigned int grid = (unsigned int)(blocks < EW_BLOCKS ? blocks : EW_BLOCKS);
  if (0 >= M || 0 >= N) return 0;
  if (3 == dtype) oz_scale_c<double><<<grid, 256, 0, (hipStream_t)hip_stream>>>((double*)c, ldc, M, N, beta);
  else if (1 == dtype) oz_scale_c<float><<<grid, 256, 0, (hipStream_t)hip_stream>>>((float*)c, ldc, M, N, beta);
  else return LXS_ENOVARIANT;
  return (int)hipGetLastError();
}

/* native FP64: C = alpha op(A) op(B) + beta C; op(A)(i, k) at a[i * ars + k * aks], op(B)(k, j) at b[k * bks + j * bcs] */
extern "C" int lxs_dgemm(const double* a, long ars, long aks, const double* b, long bks, long bcs, int M, int N, int K,
  double alpha, double beta, double* c, int ldc, void* hip_stream) {
  if (!((1 == ars) || (1 == aks)) || !((1 == bks) || (1 == bcs))) return LXS_ELIMIT;
  if (0 >= M || 0 >= N) return 0;
  if (0 >= K) return lxs_oz_scale(3, c, ldc, M, N, beta, hip_stream);
  dgemm_f64<<<dim3((unsigned int)((M + GT - 1) / GT), (unsigned int)((N + GT - 1) / GT)), 256, 0,
    (hipStream_t)hip_stream>>>(a, ars, aks, b, bks, bcs, M, N, K, alpha, beta, c, ldc);
  return (int)hipGetLastError();
}

extern "C" int lxs_oz_csplit(int dtype, const void* z, int ld, int rows, int cols, int conj, void* re, void* im,
  void* sum, void* hip_stream) {
  const long blocks = ((long)rows * cols + 255) / 256;
  const unsigned int grid = (unsigned int)(blocks < EW_BLOCKS ? blocks : EW_BLOCKS);
  if (0 >= rows || 0 >= cols) return 0;
  if (3 == dtype) {
    oz_csplit<double><<<grid, 256, 0, (hipStream_t)hip_stream>>>((const double*)z, ld, rows, cols, conj, (double*)re,
      (double*)im, (double*)sum);
  }
  else if (1 == dtype) {
    oz_csplit<float><<<grid, 256, 0, (hipStream_t)hip_stream>>>((const float*)z, ld, rows, cols, conj, (float*)re, (float*)im,
      (float*)sum);
  }
  else {
    return LXS_ENOVARIANT;
  }
  return (int)hipGetLastError();
}

extern "C" int lxs_oz_ccombine(int dtype, const void* t1, const void* t2, const void* t3, int M, int N,
  const double* alpha, const double* beta, void* c, int ldc, void* hip_stream) {
  const long blocks = ((long)M * N + 255) / 256;
  const unsigned int grid = (unsigned int)(blocks < EW_BLOCKS ? blocks : EW_BLOCKS);
  if (0 >= M || 0 >= N) return 0;
  if (3 == dtype) {
    oz_ccombine<double><<<grid, 256, 0, (hipStream_t)hip_stream>>>((const double*)t1, (const double*)t2,
      (const double*)t3, M, N, alpha[0], alpha[1], beta[0], beta[1], (double*)c, ldc);
  }
  else if (1 == dtype) {
    oz_ccombine<float><<<grid, 256, 0, (hipStream_t)hip_stream>>>((const float*)t1, (const float*)t2, (const float*)t3,
      M, N, alpha[0], alpha[1], beta[0], beta[1], (float*)c, ldc);
  }
  else {
    return LXS_ENOVARIANT;
  }
  return (int)hipGetLastError();
}
