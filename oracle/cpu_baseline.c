/*
 * ORACLE: test infrastructure only (see stack_oracle.c). CPU baseline for bench.py.
 *
 * oracle_cpu_stack times the reference host path of samples/smm/acc_bench.c:357-394:
 * libxs_gemm_index_task semantics (alpha = beta = 1, N/N, lda=m, ldb=k, ldc=m, index base 1,
 * 12-byte stack stride) with OpenMP threads that take whole C-runs, as the reference's host
 * loop does per thread (acc_bench.c:380-384). Compiled with FMA contraction and AVX2 clones.
 */
#include <stdlib.h>
#if defined(_OPENMP)
#  include <omp.h>
#endif

/* Reference host path (CPU baseline): B k-by-n column-major, threads own whole C-runs. Returns
 * the wall time in seconds of `nrepeat` passes (after `warmup` untimed passes). */
__attribute__((target_clones("avx2", "default"))) static void cpu_entry_f64(int m, int n, int k, const double* a, const double* b, double* c) {
  int j, l, i;
  for (j = 0; j < n; ++j) {
    for (l = 0; l < k; ++l) {
      const double y = b[(long)j * k + l];
      const double* const al = a + (long)l * m;
      double* const cj = c + (long)j * m;
      for (i = 0; i < m; ++i) cj[i] += al[i] * y;
    }
  }
}

__attribute__((target_clones("avx2", "default"))) static void cpu_entry_f32(int m, int n, int k, const float* a, const float* b, float* c) {
  int j, l, i;
  for (j = 0; j < n; ++j) {
    for (l = 0; l < k; ++l) {
      const float y = b[(long)j * k + l];
      const float* const al = a + (long)l * m;
      float* const cj = c + (long)j * m;
      for (i = 0; i < m; ++i) cj[i] += al[i] * y;
    }
  }
}

double oracle_cpu_stack(int type, int m, int n, int k, int stack_size, const int* stack, const void* a, const void* b,
  void* c, int nrepeat, int warmup, int nthreads) {
  int* runs = (int*)malloc(sizeof(int) * ((size_t)stack_size + 1));
  int nruns = 0, s, rep;
  double t0 = 0, t1 = 0;
  if (NULL == runs) return -1;
  for (s = 0; s < stack_size; ++s) {
    if (0 == s || stack[3 * s + 2] != stack[3 * s - 1]) runs[nruns++] = s;
  }
  runs[nruns] = stack_size;
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
          cpu_entry_f64(m, n, k, (const double*)a + a0, (const double*)b + b0, (double*)c + c0);
        }
        else {
          cpu_entry_f32(m, n, k, (const float*)a + a0, (const float*)b + b0, (float*)c + c0);
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

