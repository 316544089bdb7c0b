/*
 * ORACLE: test infrastructure only. It checks the HIP product path; it is never the product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * CPU restatement of Ozaki scheme I (int8 mantissa slicing) as the reference describes it
 * (samples/ozaki/kernels/ozaki1_int8.cl:432-589 preprocessing, :609-900 pair loop,
 * ozaki_common.cl:789-825 digit extraction; pair set samples/ozaki/ozaki_opencl.h:60-73 with cutoff
 * 2(nslices-1) - oztrim, samples/ozaki/ozaki_gemm.c:592):
 *   oracle_oz_digits_ref  the reference's digit formulation for one row: biased exponents, the
 *                         mantissa with its implicit bit aligned to the row maximum, 7-bit fields
 *                         [52-7p .. 46-7p] (the last field narrower), sign applied.
 *   oracle_oz_digits      the formulation the device uses: x = v / 2^ex (|x| < 1), then
 *                         d = trunc(128 x), x = 128 x - d, nslices times. Equal to the reference's
 *                         digits for p < 7; for p = 7 the reference keeps the 4 mantissa bits that
 *                         remain and this formulation keeps 7 bits (the next 3 come from elements
 *                         below the row maximum), so |d_7| >> 3 equals the reference's digit.
 *   oracle_ozaki_gemm     the whole emulated GEMM with the device's operation order: for every
 *                         level L = cutoff..0 the exact integer sum S_L over the pairs (p, L-p) whose
 *                         planes hold a nonzero digit, cd += 2^(-7(L+2)) * S_L in FP64, then
 *                         C = fma(alpha, ldexp(cd, ea + eb), beta C) (alpha * v if beta = 0).
 *                         Its results are therefore bit-identical to the HIP path's.
 * Zero, subnormal and non-finite inputs count as zero (ozaki_common.cl:784-805).
 *   oracle_ozaki2_gemm    Ozaki scheme II (CRT): the exact integer product of the aligned mantissas
 *                         (no residues), rounded to FP64 once and scaled (below).
 *
 * Pinning (tests/test_oracle_ozaki.py, tests/test_oracle_ozaki2.py): exact products of small-integer matrices, the rigorous
 * truncation bound against the double-double GEMM, the reference digit formulation, and planted
 * mistakes (the `plant` argument) that the pins must reject.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#if defined(_OPENMP)
#  include <omp.h>
#endif

#define OZ_NONE (-(1 << 30))
#define OZ_KB 64

/* planted mistakes, used only by the tests that prove the pins catch them */
#define OZ_PLANT_NONE 0
#define OZ_PLANT_DROPLEVEL 1 /* skip level 1 */
#define OZ_PLANT_WEIGHT 2    /* weight 2^(-7(L+1)) instead of 2^(-7(L+2)) */
#define OZ_PLANT_ROUND 3     /* digits by rounding to nearest instead of truncation */
#define OZ_PLANT_TRANSB 4    /* ignore transb */

/* ex with |v| < 2^ex for a normal finite v, OZ_NONE otherwise */
static int oz_expo(double v) {
  uint64_t bits;
  int be;
  memcpy(&bits, &v, 8);
  be = (int)((bits >> 52) & 0x7FF);
  return (0 == be || 0x7FF == be) ? OZ_NONE : be - 1022;
}

static int oz_expo_f(float v) {
  uint32_t bits;
  int be;
  memcpy(&bits, &v, 4);
  be = (int)((bits >> 23) & 0xFF);
  return (0 == be || 0xFF == be) ? OZ_NONE : be - 126;
}

/* digits of one value against row exponent ex; returns nothing, fills d[0..s-1] */
static void oz_digits1(double v, int ex, int s, int plant, signed char* d) {
  double x = (OZ_NONE != ex && OZ_NONE != oz_expo(v)) ? ldexp(v, -ex) : 0.0;
  int p;
  for (p = 0; p < s; ++p) {
    const double y = x * 128.0;
    double q = (OZ_PLANT_ROUND == plant ? nearbyint(y) : trunc(y));
    if (q > 127) q = 127;
    if (q < -127) q = -127;
    d[p] = (signed char)q;
    x = y - q;
  }
}

/* one row of n values: exponent and digits (out[p * n + i]) in the device formulation */
int oracle_oz_digits(int n, const double* v, int s, signed char* out, int* ex_out) {
  int i, ex = OZ_NONE;
  signed char d[16];
  if (1 > s || 16 < s) return 1;
  for (i = 0; i < n; ++i) {
    const int e = oz_expo(v[i]);
    if (e > ex) ex = e;
  }
  for (i = 0; i < n; ++i) {
    int p;
    oz_digits1(v[i], ex, s, OZ_PLANT_NONE, d);
    for (p = 0; p < s; ++p) out[(long)p * n + i] = d[p];
  }
  *ex_out = ex;
  return 0;
}

/* one row of n values in the reference's formulation (ieee_decompose, aligned mantissa, 7-bit fields) */
int oracle_oz_digits_ref(int n, const double* v, int s, signed char* out, int* max_biased_exp) {
  int i, p, maxe = 0;
  if (1 > s || 16 < s) return 1;
  for (i = 0; i < n; ++i) {
    uint64_t bits;
    int e;
    memcpy(&bits, &v[i], 8);
    e = (int)((bits >> 52) & 0x7FF);
    if (0x7FF != e && e > maxe) maxe = e;
  }
  for (i = 0; i < n; ++i) {
    uint64_t bits, mant, aligned;
    int e, sign, shift;
    memcpy(&bits, &v[i], 8);
    sign = (int)(bits >> 63);
    e = (int)((bits >> 52) & 0x7FF);
    mant = (bits & 0x000FFFFFFFFFFFFFull) | 0x0010000000000000ull;
    if (0 == e || 0x7FF == e) mant = 0, e = 0;
    shift = maxe - e;
    aligned = (0 != mant && shift <= 52) ? (mant >> shift) : 0;
    for (p = 0; p < s; ++p) {
      const int high = 52 - 7 * p, low = (high - 6 > 0 ? high - 6 : 0), width = high - low + 1;
      int digit = 0;
      if (width > 0 && high >= 0) digit = (int)((aligned >> low) & ((1ull << width) - 1ull));
      out[(long)p * n + i] = (signed char)(sign ? -digit : digit);
    }
  }
  *max_biased_exp = maxe;
  return 0;
}

/*
 * C = alpha op(A) op(B) + beta C emulated with s slices (column-major; use_double selects double
 * or float for A, B and C). trim: levels dropped (cutoff 2(s-1) - trim). maxk: K per preprocessing
 * pass (0: all). c is updated in place.
 */
int oracle_ozaki_gemm(int use_double, char transa, char transb, int M, int N, int K, double alpha, const void* a, int lda,
  const void* b, int ldb, double beta, void* c, int ldc, int s, int trim, int maxk, int plant) {
  const int ta = ('N' == transa || 'n' == transa) ? 0 : 1;
  const int tb = (OZ_PLANT_TRANSB == plant || 'N' == transb || 'n' == transb) ? 0 : 1;
  const int cutoff = 2 * (s - 1) - trim;
  const int kpass = (0 < maxk && maxk < K) ? maxk : K;
  int k0;
  if (0 > M || 0 > N || 0 > K || 1 > s || 8 < s || 0 > cutoff) return 1;
  if (0 == M || 0 == N) return 0;
  if (0 == K || 0 == alpha) {
    long x;
    for (x = 0; x < (long)M * N; ++x) {
      const long o = (x % M) + (x / M) * (long)ldc;
      if (use_double) ((double*)c)[o] = (0 != beta ? beta * ((double*)c)[o] : 0.0);
      else ((float*)c)[o] = (float)(0 != beta ? beta * (double)((float*)c)[o] : 0.0);
    }
    return 0;
  }
  for (k0 = 0; k0 < K; k0 += kpass) {
    const int kk = (K - k0 < kpass ? K - k0 : kpass);
    const int nkb = (kk + OZ_KB - 1) / OZ_KB, Kp = nkb * OZ_KB;
    const double beta_eff = (0 == k0 ? beta : 1.0);
    signed char* const da = (signed char*)calloc((size_t)s * M * Kp, 1);
    signed char* const db = (signed char*)calloc((size_t)s * N * Kp, 1);
    int* const ea = (int*)malloc(sizeof(int) * M);
    int* const eb = (int*)malloc(sizeof(int) * N);
    unsigned int oa = 0, ob = 0;
    int i, j;
    if (NULL == da || NULL == db || NULL == ea || NULL == eb) {
      free(da), free(db), free(ea), free(eb);
      return 2;
    }
    /* preprocessing: op(A) rows, op(B) columns */
    for (i = 0; i < M + N; ++i) {
      const int isa = (i < M), r = (isa ? i : i - M);
      const void* const src = (isa ? a : b);
      const int ld = (isa ? lda : ldb), tr = (isa ? ta : tb);
      signed char* const dst = (isa ? da : db);
      const int R = (isa ? M : N);
      int e = OZ_NONE, k, p;
      signed char d[8];
      for (k = 0; k < kk; ++k) {
        /* A: op(A)(r, k); B: op(B)(k, r) */
        const long idx = isa ? (0 == tr ? r + (long)(k0 + k) * ld : (k0 + k) + (long)r * ld)
                             : (0 == tr ? (k0 + k) + (long)r * ld : r + (long)(k0 + k) * ld);
        const double v = use_double ? ((const double*)src)[idx] : (double)((const float*)src)[idx];
        const int ev = use_double ? oz_expo(v) : oz_expo_f(((const float*)src)[idx]);
        if (ev > e) e = ev;
      }
      (isa ? ea : eb)[r] = e;
      for (k = 0; k < kk; ++k) {
        const long idx = isa ? (0 == tr ? r + (long)(k0 + k) * ld : (k0 + k) + (long)r * ld)
                             : (0 == tr ? (k0 + k) + (long)r * ld : r + (long)(k0 + k) * ld);
        const double v = use_double ? ((const double*)src)[idx] : (double)((const float*)src)[idx];
        const int valid = use_double ? (OZ_NONE != oz_expo(v)) : (OZ_NONE != oz_expo_f(((const float*)src)[idx]));
        oz_digits1(valid ? v : 0.0, e, s, plant, d);
        for (p = 0; p < s; ++p) {
          dst[((long)p * R + r) * Kp + k] = d[p];
          if (0 != d[p]) {
            if (isa) oa |= 1u << p;
            else ob |= 1u << p;
          }
        }
      }
    }
#if defined(_OPENMP)
#  pragma omp parallel for schedule(dynamic, 1) private(i)
#endif
    for (j = 0; j < N; ++j) {
      for (i = 0; i < M; ++i) {
        double cd = 0, v, cv;
        int L;
        for (L = cutoff; L >= 0; --L) {
          const double scale = ldexp(1.0, -7 * (L + (OZ_PLANT_WEIGHT == plant ? 1 : 2)));
          int64_t acc = 0;
          int p;
          if (OZ_PLANT_DROPLEVEL == plant && 1 == L) continue;
          for (p = (L - s + 1 > 0 ? L - s + 1 : 0); p <= (L < s - 1 ? L : s - 1); ++p) {
            const signed char* ra;
            const signed char* rb;
            int kb;
            if (!((oa >> p) & (ob >> (L - p)) & 1u)) continue;
            ra = da + ((long)p * M + i) * Kp;
            rb = db + ((long)(L - p) * N + j) * Kp;
            for (kb = 0; kb < nkb; ++kb) {
              int32_t blk = 0;
              int k;
              for (k = kb * OZ_KB; k < kb * OZ_KB + OZ_KB; ++k) blk += (int32_t)ra[k] * (int32_t)rb[k];
              acc += blk;
            }
          }
          cd += scale * (double)acc;
        }
        v = (OZ_NONE != ea[i] && OZ_NONE != eb[j]) ? ldexp(cd, ea[i] + eb[j]) : 0.0;
        if (use_double) {
          double* const pc = (double*)c + i + (long)j * ldc;
          cv = (0 != beta_eff ? fma(alpha, v, beta_eff * *pc) : alpha * v);
          *pc = cv;
        }
        else {
          float* const pc = (float*)c + i + (long)j * ldc;
          cv = (0 != beta_eff ? fma(alpha, v, beta_eff * (double)*pc) : alpha * v);
          *pc = (float)cv;
        }
      }
    }
    free(da), free(db), free(ea), free(eb);
  }
  return 0;
}

/*
 * Ozaki scheme II (CRT), the integer product the reference recovers from its residues
 * (samples/ozaki/kernels/ozaki2_int8.cl:1808-1860 preprocessing: per-row (per-column for B) maximum
 * biased exponent over normal finite values, the mantissa with its implicit bit shifted right by the
 * exponent difference plus the trimmed bits, sign kept; :1329-1440 Garner and Horner: the exact sum
 * of products, scaled by 2^(ea + eb - 2 (bias + mantissa bits - trim))). The oracle forms the same
 * integers and their exact sum (128-bit) directly, without residues, and rounds it to FP64 once; then
 * C = fma(alpha, x 2^sh, beta C) (alpha x 2^sh if beta = 0), per K pass of kpass (the first pass
 * applies beta, later ones accumulate). float C receives that double rounded to float.
 * Planted mistakes (tests only): 1 rounds the aligned mantissa to nearest, 2 drops the sign of B,
 * 3 scales by 2^(sh + 1), 4 sums the products in FP64.
 */
#define OZ2_PLANT_ROUND 1
#define OZ2_PLANT_SIGNB 2
#define OZ2_PLANT_SCALE 3
#define OZ2_PLANT_FPSUM 4

/* biased exponent and mantissa with its implicit bit (0, 0 for zero, subnormal and non-finite) */
static void oz2_decompose(int use_double, const void* src, long idx, int* sign, int* be, uint64_t* mant) {
  if (use_double) {
    uint64_t bits;
    memcpy(&bits, (const double*)src + idx, 8);
    *sign = (int)(bits >> 63);
    *be = (int)((bits >> 52) & 0x7FF);
    *mant = (bits & 0x000FFFFFFFFFFFFFull) | 0x0010000000000000ull;
    if (0 == *be || 0x7FF == *be) *be = 0, *mant = 0;
  }
  else {
    uint32_t bits;
    memcpy(&bits, (const float*)src + idx, 4);
    *sign = (int)(bits >> 31);
    *be = (int)((bits >> 23) & 0xFF);
    *mant = (uint64_t)((bits & 0x007FFFFFu) | 0x00800000u);
    if (0 == *be || 0xFF == *be) *be = 0, *mant = 0;
  }
}

int oracle_ozaki2_gemm(int use_double, char transa, char transb, int M, int N, int K, double alpha, const void* a,
  int lda, const void* b, int ldb, double beta, void* c, int ldc, int trim, int kpass, int plant) {
  const int ta = ('N' == transa || 'n' == transa) ? 0 : 1;
  const int tb = ('N' == transb || 'n' == transb) ? 0 : 1;
  const int mant_bits = (use_double ? 52 : 23), bias_plus_mant = (use_double ? 1075 : 150);
  int k0;
  if (0 > M || 0 > N || 0 > K || 0 > trim || mant_bits < trim) return 1;
  if (0 >= kpass || K < kpass) kpass = (0 < K ? K : 1);
  if (0 == M || 0 == N) return 0;
  if (0 == K || 0 == alpha) {
    long x;
    for (x = 0; x < (long)M * N; ++x) {
      const long o = (x % M) + (x / M) * (long)ldc;
      if (use_double) ((double*)c)[o] = (0 != beta ? beta * ((double*)c)[o] : 0.0);
      else ((float*)c)[o] = (float)(0 != beta ? beta * (double)((float*)c)[o] : 0.0);
    }
    return 0;
  }
  for (k0 = 0; k0 < K; k0 += kpass) {
    const int kk = (K - k0 < kpass ? K - k0 : kpass);
    const double beta_eff = (0 == k0 ? beta : 1.0);
    int64_t* const ia = (int64_t*)malloc(sizeof(int64_t) * (size_t)M * kk);
    int64_t* const ib = (int64_t*)malloc(sizeof(int64_t) * (size_t)N * kk);
    int* const ea = (int*)malloc(sizeof(int) * M);
    int* const eb = (int*)malloc(sizeof(int) * N);
    int i, j;
    if (NULL == ia || NULL == ib || NULL == ea || NULL == eb) {
      free(ia), free(ib), free(ea), free(eb);
      return 2;
    }
    for (i = 0; i < M + N; ++i) { /* op(A) rows, then op(B) columns */
      const int isa = (i < M), r = (isa ? i : i - M), tr = (isa ? ta : tb), ld = (isa ? lda : ldb);
      const void* const src = (isa ? a : b);
      int64_t* const dst = (isa ? ia : ib) + (long)r * kk;
      int k, maxe = 0;
      for (k = 0; k < kk; ++k) {
        const long idx = isa ? (0 == tr ? r + (long)(k0 + k) * ld : (k0 + k) + (long)r * ld)
                             : (0 == tr ? (k0 + k) + (long)r * ld : r + (long)(k0 + k) * ld);
        int sign, be;
        uint64_t mant;
        oz2_decompose(use_double, src, idx, &sign, &be, &mant);
        if (be > maxe) maxe = be;
      }
      (isa ? ea : eb)[r] = maxe;
      for (k = 0; k < kk; ++k) {
        const long idx = isa ? (0 == tr ? r + (long)(k0 + k) * ld : (k0 + k) + (long)r * ld)
                             : (0 == tr ? (k0 + k) + (long)r * ld : r + (long)(k0 + k) * ld);
        int sign, be;
        uint64_t mant, aligned = 0;
        oz2_decompose(use_double, src, idx, &sign, &be, &mant);
        if (0 != mant) {
          const int shift = (maxe - be) + trim;
          if (shift <= mant_bits) {
            aligned = mant >> shift;
            if (OZ2_PLANT_ROUND == plant && 0 < shift && 0 != ((mant >> (shift - 1)) & 1)) ++aligned;
          }
        }
        if (OZ2_PLANT_SIGNB == plant && !isa) sign = 0;
        dst[k] = (sign ? -(int64_t)aligned : (int64_t)aligned);
      }
    }
#if defined(_OPENMP)
#  pragma omp parallel for schedule(dynamic, 1) private(i)
#endif
    for (j = 0; j < N; ++j) {
      for (i = 0; i < M; ++i) {
        const int64_t* const ra = ia + (long)i * kk;
        const int64_t* const rb = ib + (long)j * kk;
        const int sh = ea[i] + eb[j] - 2 * (bias_plus_mant - trim) + (OZ2_PLANT_SCALE == plant ? 1 : 0);
        const long o = i + (long)j * ldc;
        double v, val;
        int k;
        if (OZ2_PLANT_FPSUM == plant) {
          v = 0;
          for (k = 0; k < kk; ++k) v += (double)ra[k] * (double)rb[k];
        }
        else {
          __int128 x = 0;
          for (k = 0; k < kk; ++k) x += (__int128)ra[k] * (__int128)rb[k];
          v = (double)x; /* round to nearest even, once */
        }
        val = ldexp(v, sh);
        if (use_double) {
          double* const pc = (double*)c + o;
          *pc = (0 != beta_eff ? fma(alpha, val, beta_eff * *pc) : alpha * val);
        }
        else {
          float* const pc = (float*)c + o;
          *pc = (float)(0 != beta_eff ? fma(alpha, val, beta_eff * (double)*pc) : alpha * val);
        }
      }
    }
    free(ia), free(ib), free(ea), free(eb);
  }
  return 0;
}
