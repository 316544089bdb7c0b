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
#include "lxs_kernels.h"

namespace {

typedef int i4_t __attribute__((ext_vector_type(4)));
typedef int i16_t __attribute__((ext_vector_type(16)));
typedef double d4_t __attribute__((ext_vector_type(4)));

constexpr int OZ_TILE = 64; /* rows x k of a preprocessing tile */
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

/* load a 64x64 tile (rows r0.., k k0..) into LDS t[r][k] as double, coalesced along the contiguous
   dimension; elements outside R x K, zero, subnormal or non-finite become 0 */
template <typename T>
__device__ __forceinline__ void oz_load_tile(double (*t)[OZ_TILE + 1], const T* __restrict__ p, long rs, long ks,
  int r0, int k0, int R, int K) {
  const int tid = threadIdx.x;
  if (1 == ks) { /* k contiguous: 16 threads per row, 4 consecutive k each; 16 rows per pass */
#pragma unroll
    for (int pass = 0; pass < 4; ++pass) {
      const int r = pass * 16 + (tid >> 4), kq = (tid & 15) * 4;
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const int gr = r0 + r, gk = k0 + kq + x;
        const T v = (gr < R && gk < K) ? p[(long)gr * rs + gk] : T(0);
        t[r][kq + x] = (OZ_EXP_NONE != oz_expo(v)) ? (double)v : 0.0;
      }
    }
  }
  else { /* rows contiguous: 16 threads per k, 4 consecutive rows each; 16 k per pass */
#pragma unroll
    for (int pass = 0; pass < 4; ++pass) {
      const int k = pass * 16 + (tid >> 4), rq = (tid & 15) * 4;
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const int gr = r0 + rq + x, gk = k0 + k;
        const T v = (gr < R && gk < K) ? p[gr + (long)gk * ks] : T(0);
        t[rq + x][k] = (OZ_EXP_NONE != oz_expo(v)) ? (double)v : 0.0;
      }
    }
  }
}

__global__ void __launch_bounds__(256) oz_init(int* __restrict__ e, int n, int* __restrict__ occ) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) e[i] = OZ_EXP_NONE;
  if (0 == blockIdx.x && (int)threadIdx.x < 8) occ[threadIdx.x] = 0;
}

template <typename T>
__global__ void __launch_bounds__(256) oz_exp(const T* __restrict__ p, long rs, long ks, int R, int K,
  int* __restrict__ e) {
  __shared__ double t[OZ_TILE][OZ_TILE + 1];
  const int r0 = blockIdx.x * OZ_TILE, k0 = blockIdx.y * OZ_TILE;
  oz_load_tile<T>(t, p, rs, ks, r0, k0, R, K);
  __syncthreads();
  { /* 4 threads per row, 16 k each, then a 4-lane maximum */
    const int r = threadIdx.x >> 2, kq = (threadIdx.x & 3) * 16;
    int ex = OZ_EXP_NONE;
#pragma unroll
    for (int x = 0; x < 16; ++x) ex = max(ex, oz_expo(t[r][kq + x]));
    ex = max(ex, __shfl_xor(ex, 1));
    ex = max(ex, __shfl_xor(ex, 2));
    if (0 == (threadIdx.x & 3) && OZ_EXP_NONE != ex && r0 + r < R) atomicMax(e + r0 + r, ex);
  }
}

/* s int8 digits of 16 consecutive k of one row, written as one 16-byte vector per plane */
template <typename T>
__global__ void __launch_bounds__(256) oz_slice(const T* __restrict__ p, long rs, long ks, int R, int K,
  const int* __restrict__ e, int s, int Rp, int Kp, int8_t* __restrict__ planes, int* __restrict__ occ) {
  __shared__ double t[OZ_TILE][OZ_TILE + 1];
  const int r0 = blockIdx.x * OZ_TILE, k0 = blockIdx.y * OZ_TILE;
  const int tid = threadIdx.x, r = tid >> 2, kq = (tid & 3) * 16;
  oz_load_tile<T>(t, p, rs, ks, r0, k0, R, K);
  __syncthreads();
  const int row = r0 + r;
  const int ex = (row < R ? e[row] : OZ_EXP_NONE);
  double x[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) x[i] = (OZ_EXP_NONE != ex ? ldexp(t[r][kq + i], -ex) : 0.0); /* |x| < 1, exact */
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
    *(i4_t*)(planes + ((long)pl * Rp + row) * Kp + k0 + kq) = d.v;
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
constexpr int LT = 256;        /* int8 level GEMM: output tile (i and j) per work-group */
constexpr int LK = 128;        /* k padding of the digit planes */

/*
 * One work item: level L (pairs p = plo .. phi, q = L - p), units u0 .. u1 of the level (unit u is
 * pair plo + u / nkb, k-block u % nkb of 64 k). Every item writes its own int32 partial sums, exact
 * because an item holds at most 2080 units: 2080 * 64 * 127^2 < 2^31.
 */
struct oz_item_t {
  int level, u0, u1, pad;
};
constexpr int OZ_ITEMS = 128; /* items per launch, passed by value */
struct oz_items_t {
  oz_item_t it[OZ_ITEMS];
};

/*
 * C32[z](j, i) = sum over the item's units of sum_k A_p(i, k) B_q(j, k), stored at z * Mp * Np + j * Mp + i.
 * v_mfma_i32_32x32x32_i8 with the B planes as the first operand (rows j) and the A planes as the
 * second (columns i): lane l of D holds column i = l % 32 and rows 8(r/4) + 4(l/32) + r%4, so for a
 * fixed register the 32 lanes of a half-wave store 128 contiguous bytes along i.
 * 4 wavefronts in 2 x 2, each 128 x 128 (4 x 4 MFMA tiles, 256 accumulator registers).
 * Operand tiles of 64 k (one "unit") move from global memory straight into LDS by DMA
 * (global_load_lds_dwordx4), four stages deep (three in flight while one is computed), without
 * staging registers.
 * LDS row layout: 64 bytes per row (4 chunks of 16), chunk c of row r stored at position
 * c ^ ((r >> 2) & 3); the 16 lanes of a quarter-wave read rows r..r+15 of one chunk from 16 distinct
 * 4-bank groups. One DMA instruction fills 16 rows (1 KB): lane l writes position l % 4 of row l / 4,
 * so it loads chunk (l % 4) ^ ((row >> 2) & 3) of that row.
 */
constexpr int QK = 64;                /* k per stage */
constexpr int QTILE = LT * QK;        /* bytes of one operand tile per stage (16 KB) */

/* operand tile of zeros: the source of the padding stages (their MFMAs add nothing) */
__device__ __attribute__((aligned(1024))) int8_t oz_zero_tile[QTILE];

__device__ __forceinline__ void oz_dma16(const int8_t* src, int8_t* lds_dst) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
    (__attribute__((address_space(3))) void*)lds_dst, 16, 0, 0);
}

__global__ void __launch_bounds__(256) oz_level_gemm(const int8_t* __restrict__ pa, const int8_t* __restrict__ pb, const int* __restrict__ occa,
  const int* __restrict__ occb, oz_items_t items, int s, int Mp, int Np, int Kp, int* __restrict__ c32) {
  /* one array per stage ([A|B][256 rows x 64 bytes]): distinct objects, so the compiler can tell the DMA
     targets of the stages in flight from the stage being read */
  __shared__ __attribute__((aligned(1024))) int8_t ls0[2 * QTILE], ls1[2 * QTILE], ls2[2 * QTILE], ls3[2 * QTILE];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wj = wave >> 1, wi = wave & 1;
  const int i0 = blockIdx.x * LT, j0 = blockIdx.y * LT;
  const oz_item_t it = items.it[blockIdx.z];
  const int L = it.level, nkb = Kp / QK;
  const int plo = (L - s + 1 > 0 ? L - s + 1 : 0);
  unsigned int oa = 0, ob = 0;
  for (int p = 0; p < s; ++p) {
    oa |= (0 != occa[p] ? 1u : 0u) << p;
    ob |= (0 != occb[p] ? 1u : 0u) << p;
  }
  oa = __builtin_amdgcn_readfirstlane(oa);
  ob = __builtin_amdgcn_readfirstlane(ob);
  i16_t acc[4][4];
#pragma unroll
  for (int x = 0; x < 4; ++x) {
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = i16_t{0};
  }
  /* active units (pairs with both planes occupied) of this item, counted up front */
  int nact = 0;
  for (int pp = it.u0 / nkb; pp * nkb < it.u1; ++pp) {
    const int p = plo + pp, lo = (it.u0 > pp * nkb ? it.u0 : pp * nkb), hi = (it.u1 < (pp + 1) * nkb ? it.u1 : (pp + 1) * nkb);
    if ((oa >> p) & (ob >> (L - p)) & 1u) nact += hi - lo;
  }
  if (0 < nact) {
    int fu = it.u0; /* next unit to load */
    auto skip = [&]() {
      while (fu < it.u1 && !((oa >> (plo + fu / nkb)) & (ob >> (L - plo - fu / nkb)) & 1u)) fu = (fu / nkb + 1) * nkb;
    };
    skip();
    const int drow = lane >> 2, dpos = lane & 3; /* this lane's row of a DMA instruction's 16, and position */
    /* stage g into stage array buf: A tile rows 16 q .. 16 q + 15 for q = wave + 4 r, B alike; the next
       active unit while g < nact, the zero tile after that (the stage count is padded to a multiple of 4) */
    int g = 0;
    auto dma = [&](int8_t* buf) {
      const int p = plo + fu / nkb, kb = fu - (fu / nkb) * nkb;
      const bool real = (g < nact);
      const int8_t* const ga = (real ? pa + ((long)p * Mp + i0) * Kp + kb * QK : oz_zero_tile);
      const int8_t* const gb = (real ? pb + ((long)(L - p) * Np + j0) * Kp + kb * QK : oz_zero_tile);
      const long stride = (real ? Kp : QK);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int q = wave + 4 * r, row = 16 * q + drow, chunk = dpos ^ ((row >> 2) & 3);
        oz_dma16(ga + (long)row * stride + 16 * chunk, buf + q * 1024);
        oz_dma16(gb + (long)row * stride + 16 * chunk, buf + QTILE + q * 1024);
      }
      if (real) {
        ++fu;
        skip();
      }
      ++g;
    };
    const int frow = lane & 31, fswz = (frow >> 2) & 3; /* (row >> 2) & 3 is the same for rows 32 x + frow */
    const int fpos0 = (lane >> 5) ^ fswz, fpos1 = (2 + (lane >> 5)) ^ fswz;
    /* one stage: wait for this wave's loads of the stage (two later stages may fly), barrier (every wave's
       loads of the stage have landed, every wave has consumed the previous stage), refill the array
       computed last time, then all 16 fragments and 32 MFMAs */
#define OZ_STAGE(CUR, NXT) \
  do { \
    __builtin_amdgcn_s_waitcnt(0x4F70); /* vmcnt(16): this wave's DMA of the stage has landed */ \
    __asm__ __volatile__("" ::: "memory"); \
    __builtin_amdgcn_s_barrier(); /* no fence: it would wait for the stages still in flight too */ \
    __asm__ __volatile__("" ::: "memory"); \
    dma(NXT); \
    const int8_t* const la_ = CUR + (128 * wi + frow) * QK; \
    const int8_t* const lb_ = CUR + QTILE + (128 * wj + frow) * QK; \
    i4_t fa0_[4], fb0_[4], fa1_[4], fb1_[4]; \
    _Pragma("unroll") for (int x_ = 0; x_ < 4; ++x_) { \
      fb0_[x_] = *(const i4_t*)(lb_ + 32 * x_ * QK + 16 * fpos0); \
      fa0_[x_] = *(const i4_t*)(la_ + 32 * x_ * QK + 16 * fpos0); \
      fb1_[x_] = *(const i4_t*)(lb_ + 32 * x_ * QK + 16 * fpos1); \
      fa1_[x_] = *(const i4_t*)(la_ + 32 * x_ * QK + 16 * fpos1); \
    } \
    _Pragma("unroll") for (int x_ = 0; x_ < 4; ++x_) { \
      _Pragma("unroll") for (int y_ = 0; y_ < 4; ++y_) { \
        acc[x_][y_] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fb0_[x_], fa0_[y_], acc[x_][y_], 0, 0, 0); \
      } \
    } \
    _Pragma("unroll") for (int x_ = 0; x_ < 4; ++x_) { \
      _Pragma("unroll") for (int y_ = 0; y_ < 4; ++y_) { \
        acc[x_][y_] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fb1_[x_], fa1_[y_], acc[x_][y_], 0, 0, 0); \
      } \
    } \
  } while (0)
    dma(ls0);
    dma(ls1);
    dma(ls2);
    for (int t = 0; t < nact; t += 4) { /* one exit: the accumulators stay in their registers */
      OZ_STAGE(ls0, ls3);
      OZ_STAGE(ls1, ls0);
      OZ_STAGE(ls2, ls1);
      OZ_STAGE(ls3, ls2);
    }
#undef OZ_STAGE
    __builtin_amdgcn_s_waitcnt(0); /* the trailing zero-tile loads land before the work-group ends */
  }
  int* const out = c32 + (long)blockIdx.z * Mp * Np;
#pragma unroll
  for (int x = 0; x < 4; ++x) {
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int i = i0 + 128 * wi + 32 * y + (lane & 31);
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int j = j0 + 128 * wj + 32 * x + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        out[(long)j * Mp + i] = acc[x][y][r];
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
  const dim3 grid((unsigned int)(Rp / OZ_TILE), (unsigned int)(Kp / OZ_TILE));
  if (0 != Rp % LT || 0 != Kp % LK || Rp < R || Kp < K || 0 >= Kp || 1 > s || 8 < s) return LXS_ELIMIT;
  if (!((1 == rs) || (1 == ks))) return LXS_ELIMIT;
  oz_init<<<(Rp + 255) / 256, 256, 0, st>>>(e, Rp, occ);
  if (3 == dtype) {
    oz_exp<double><<<grid, 256, 0, st>>>((const double*)p, rs, ks, R, K, e);
    oz_slice<double><<<grid, 256, 0, st>>>((const double*)p, rs, ks, R, K, e, s, Rp, Kp, planes, occ);
  }
  else if (1 == dtype) {
    oz_exp<float><<<grid, 256, 0, st>>>((const float*)p, rs, ks, R, K, e);
    oz_slice<float><<<grid, 256, 0, st>>>((const float*)p, rs, ks, R, K, e, s, Rp, Kp, planes, occ);
  }
  else {
    return LXS_ENOVARIANT;
  }
  return (int)hipGetLastError();
}

/* items: host array of nitems x {level, u0, u1, pad}; partial z of item z at c32 + z * Mp * Np;
   planes row-padded to a multiple of 256 and k-padded to a multiple of 128 */
extern "C" int lxs_oz_level_gemm(const signed char* pa, const signed char* pb, const int* occa, const int* occb,
  const int* items, int nitems, int s, int Mp, int Np, int Kp, int* c32, void* hip_stream) {
  if (0 != Mp % LT || 0 != Np % LT || 0 != Kp % LK || 0 >= Kp || 1 > s || 8 < s || 0 >= nitems) return LXS_ELIMIT;
  for (int z = 0; z < nitems; ++z) { /* units of 64 k within the level's pairs, at most 2080 per item */
    const int L = items[4 * z], plo = (L - s + 1 > 0 ? L - s + 1 : 0), phi = (L < s - 1 ? L : s - 1);
    if (0 > L || 2 * (s - 1) < L || 0 > items[4 * z + 1] || items[4 * z + 1] >= items[4 * z + 2] ||
        2080 < items[4 * z + 2] - items[4 * z + 1] || (phi - plo + 1) * (Kp / QK) < items[4 * z + 2]) {
      return LXS_ELIMIT;
    }
  }
  for (int z0 = 0; z0 < nitems; z0 += OZ_ITEMS) {
    const int nz = (nitems - z0 < OZ_ITEMS ? nitems - z0 : OZ_ITEMS);
    oz_items_t arg;
    for (int z = 0; z < OZ_ITEMS; ++z) {
      const int* const src = items + 4 * (z0 + (z < nz ? z : 0));
      arg.it[z].level = src[0];
      arg.it[z].u0 = src[1];
      arg.it[z].u1 = src[2];
      arg.it[z].pad = 0;
    }
    oz_level_gemm<<<dim3((unsigned int)(Mp / LT), (unsigned int)(Np / LT), (unsigned int)nz), 256, 0,
      (hipStream_t)hip_stream>>>(pa, pb, occa, occb, arg, s, Mp, Np, Kp, c32 + (long)z0 * Mp * Np);
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
