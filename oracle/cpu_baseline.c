/*
 * ORACLE: test infrastructure only (see stack_oracle.c). CPU baseline for bench.py.
 *
 * oracle_cpu_stack times the reference host path of samples/smm/acc_bench.c:357-394:
 * libxs_gemm_index_task semantics (alpha = beta = 1, N/N, lda=m, ldb=k, ldc=m, index base 1,
 * 12-byte stack stride) with OpenMP threads that take whole C-runs, as the reference's host
 * loop does per thread (acc_bench.c:380-384). Compiled with FMA contraction, in AVX-512 and AVX2
 * versions chosen at run time.
 */
#include <immintrin.h>
#include <stdlib.h>
#if defined(_OPENMP)
#  include <omp.h>
#endif

/* One stack entry of the reference host path, C(m x n) += A(m x k) B(k x n), all column-major (B
 * untransposed, ldb = k). Column blocks of W rows of C stay in registers over the whole k loop, four
 * columns at a time, so each loaded A vector feeds four independent FMA chains; rows past the last
 * whole block take a plain loop. Compiled three times (AVX-512 + FMA, AVX2 + FMA, baseline x86-64) and
 * chosen at run time by the CPU's features. */
#define CPU_ENTRY(NAME, T, W, ATTR) \
  ATTR static void NAME(int m, int n, int k, const T* restrict a, const T* restrict b, T* restrict c) { \
    int j = 0, i0, l, i, q; \
    for (; j + 4 <= n; j += 4) { \
      for (i0 = 0; i0 + W <= m; i0 += W) { \
        T acc[4][W]; \
        for (q = 0; q < 4; ++q) { \
          for (i = 0; i < W; ++i) acc[q][i] = c[(long)(j + q) * m + i0 + i]; \
        } \
        for (l = 0; l < k; ++l) { \
          const T* const al = a + (long)l * m + i0; \
          for (q = 0; q < 4; ++q) { \
            const T y = b[(long)(j + q) * k + l]; \
            for (i = 0; i < W; ++i) acc[q][i] += al[i] * y; \
          } \
        } \
        for (q = 0; q < 4; ++q) { \
          for (i = 0; i < W; ++i) c[(long)(j + q) * m + i0 + i] = acc[q][i]; \
        } \
      } \
      for (q = 0; q < 4 && i0 < m; ++q) { \
        T* const cj = c + (long)(j + q) * m; \
        for (l = 0; l < k; ++l) { \
          const T y = b[(long)(j + q) * k + l]; \
          for (i = i0; i < m; ++i) cj[i] += a[(long)l * m + i] * y; \
        } \
      } \
    } \
    for (; j < n; ++j) { \
      T* const cj = c + (long)j * m; \
      for (l = 0; l < k; ++l) { \
        const T y = b[(long)j * k + l]; \
        for (i = 0; i < m; ++i) cj[i] += a[(long)l * m + i] * y; \
      } \
    } \
  }

/* AVX-512: the same blocking with masked loads and stores for the rows past the last whole block */
#define CPU_ENTRY512(NAME, T, W, VT, MT, LOADZ, STORE, FMA, SET1, ALLMASK) \
  __attribute__((target("avx512f,avx512vl,fma"))) static void NAME(int m, int n, int k, const T* restrict a, \
    const T* restrict b, T* restrict c) { \
    int j = 0, i0, l, q; \
    for (; j < n; j += 4) { \
      const int nq = (n - j < 4 ? n - j : 4); \
      for (i0 = 0; i0 < m; i0 += W) { \
        const MT mask = (MT)(m - i0 >= W ? ALLMASK : ((1u << (m - i0)) - 1u)); \
        VT acc0 = LOADZ(mask, c + (long)j * m + i0); \
        VT acc1 = (1 < nq ? LOADZ(mask, c + (long)(j + 1) * m + i0) : acc0); \
        VT acc2 = (2 < nq ? LOADZ(mask, c + (long)(j + 2) * m + i0) : acc0); \
        VT acc3 = (3 < nq ? LOADZ(mask, c + (long)(j + 3) * m + i0) : acc0); \
        const T* const b0 = b + (long)j * k; \
        const T* const b1 = b + (long)(j + (1 < nq ? 1 : 0)) * k; \
        const T* const b2 = b + (long)(j + (2 < nq ? 2 : 0)) * k; \
        const T* const b3 = b + (long)(j + (3 < nq ? 3 : 0)) * k; \
        for (l = 0; l < k; ++l) { \
          const VT av = LOADZ(mask, a + (long)l * m + i0); \
          acc0 = FMA(av, SET1(b0[l]), acc0); \
          acc1 = FMA(av, SET1(b1[l]), acc1); \
          acc2 = FMA(av, SET1(b2[l]), acc2); \
          acc3 = FMA(av, SET1(b3[l]), acc3); \
        } \
        STORE(c + (long)j * m + i0, mask, acc0); \
        for (q = 1; q < nq; ++q) STORE(c + (long)(j + q) * m + i0, mask, 1 == q ? acc1 : (2 == q ? acc2 : acc3)); \
      } \
    } \
  }

CPU_ENTRY512(cpu_entry_f64_avx512, double, 8, __m512d, __mmask8, _mm512_maskz_loadu_pd, _mm512_mask_storeu_pd,
  _mm512_fmadd_pd, _mm512_set1_pd, 0xFFu)
CPU_ENTRY(cpu_entry_f64_avx2, double, 8, __attribute__((target("avx2,fma"))))
CPU_ENTRY(cpu_entry_f64_base, double, 8, )
CPU_ENTRY512(cpu_entry_f32_avx512, float, 16, __m512, __mmask16, _mm512_maskz_loadu_ps, _mm512_mask_storeu_ps,
  _mm512_fmadd_ps, _mm512_set1_ps, 0xFFFFu)
CPU_ENTRY(cpu_entry_f32_avx2, float, 16, __attribute__((target("avx2,fma"))))
CPU_ENTRY(cpu_entry_f32_base, float, 16, )

typedef void (*cpu_entry_f64_t)(int, int, int, const double*, const double*, double*);
typedef void (*cpu_entry_f32_t)(int, int, int, const float*, const float*, float*);

/* 2: AVX-512 + FMA, 1: AVX2 + FMA, 0: baseline */
static int cpu_isa(void) {
  __builtin_cpu_init();
  if (__builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512vl") && __builtin_cpu_supports("fma")) {
    return 2;
  }
  return (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) ? 1 : 0;
}

int oracle_cpu_isa(void) { return cpu_isa(); }

double oracle_cpu_stack(int type, int m, int n, int k, int stack_size, const int* stack, const void* a, const void* b,
  void* c, int nrepeat, int warmup, int nthreads) {
  int* runs = (int*)malloc(sizeof(int) * ((size_t)stack_size + 1));
  int nruns = 0, s, rep;
  cpu_entry_f64_t f64;
  cpu_entry_f32_t f32;
  double t0 = 0, t1 = 0;
  if (NULL == runs) return -1;
  for (s = 0; s < stack_size; ++s) {
    if (0 == s || stack[3 * s + 2] != stack[3 * s - 1]) runs[nruns++] = s;
  }
  runs[nruns] = stack_size;
  {
    const int isa = cpu_isa();
    f64 = (2 == isa ? cpu_entry_f64_avx512 : (1 == isa ? cpu_entry_f64_avx2 : cpu_entry_f64_base));
    f32 = (2 == isa ? cpu_entry_f32_avx512 : (1 == isa ? cpu_entry_f32_avx2 : cpu_entry_f32_base));
  }
#if defined(_OPENMP)
  if (0 < nthreads) omp_set_num_threads(nthreads);
#endif
  for (rep = 0; rep < warmup + nrepeat; ++rep) {
    int r;
    if (rep == warmup) {
#if defined(_OPENMP)
      t0 = omp_get_wtime();
#endif
    }
#if defined(_OPENMP)
#  pragma omp parallel for schedule(dynamic, 8)
#endif
    for (r = 0; r < nruns; ++r) {
      int e;
      for (e = runs[r]; e < runs[r + 1]; ++e) {
        const long a0 = stack[3 * e] - 1, b0 = stack[3 * e + 1] - 1, c0 = stack[3 * e + 2] - 1;
        if (3 == type) {
          f64(m, n, k, (const double*)a + a0, (const double*)b + b0, (double*)c + c0);
        }
        else {
          f32(m, n, k, (const float*)a + a0, (const float*)b + b0, (float*)c + c0);
        }
      }
    }
  }
#if defined(_OPENMP)
  t1 = omp_get_wtime();
#endif
  free(runs);
  return t1 - t0;
}

