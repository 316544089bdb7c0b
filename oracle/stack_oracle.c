/*
 * ORACLE: test infrastructure only. It checks the HIP product path; it is never the product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * CPU restatement of the reference's DBCSR stack-multiply path (hfp/libxstream):
 *   oracle_stack_dd     C[c_i] += A[a_i] * B[b_i] over a (a,b,c) stack of 1-based element
 *                       offsets, A m-by-k column-major, C m-by-n column-major, B either k-by-n
 *                       column-major (host view, samples/smm/acc_bench.c:363-374 with ldb=k) or
 *                       row-major (device view after the transpose, kernels/multiply.cl:12-15,
 *                       libxstream/opencl/libxstream_common.h:86-87). Every product is formed
 *                       exactly (TwoProd via fma, Dekker/Knuth) and summed in double-double
 *                       (TwoSum), so the result carries ~106 bits before the final rounding.
 *   oracle_error        max normalised error max|C-C_ref| / max|C_ref| and the normwise
 *                       ||C-C_ref||_F / ||C_ref||_F against the double-double result.
 *   oracle_transpose    kernels/transpose.cl:39-45: column-major m-by-n -> row-major, in place.
 *   oracle_glibc_rand   the rnd table of samples/smm/acc_bench.c:144-147 (srand, rand()).
 *   oracle_init_stack   INIT_STACK, samples/smm/acc_bench.h:48-79.
 *   oracle_init_mat     INIT_MAT, samples/smm/acc_bench.h:30-40.
 *   (the timed CPU baseline, oracle_cpu_stack, is in cpu_baseline.c)
 *
 * Pinning (tests/test_oracle.py): exact rational arithmetic on small stacks, fixtures under
 * tests/golden/, closed-form special cases, and planted mistakes (the `plant` argument below)
 * that the pins must reject.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#if defined(_OPENMP)
#  include <omp.h>
#endif

/* planted mistakes, used only by the tests that prove the pins catch them */
#define ORACLE_PLANT_NONE 0
#define ORACLE_PLANT_BLAYOUT 1   /* read B with the other layout */
#define ORACLE_PLANT_DROPK 2     /* drop the last k term */
#define ORACLE_PLANT_BASE0 3     /* treat offsets as 0-based */
#define ORACLE_PLANT_NOACC 4     /* overwrite C instead of accumulating */

static inline void two_sum(double a, double b, double* s, double* e) {
  const double x = a + b, z = x - a;
  *e = (a - (x - z)) + (b - z);
  *s = x;
}

static inline void dd_add(double* hi, double* lo, double xh, double xl) {
  double s, e, t;
  two_sum(*hi, xh, &s, &e);
  e += *lo + xl;
  t = s + e; /* fast two-sum renormalisation */
  *lo = e - (t - s);
  *hi = t;
}

static inline double load(const void* p, int type, long i) {
  return (3 == type) ? ((const double*)p)[i] : (double)((const float*)p)[i];
}

int oracle_stack_dd(int type, int m, int n, int k, int stack_size, const int* stack, const void* a, const void* b,
  int b_rowmajor, const void* c_in, long c_elems, double* c_hi, double* c_lo, int plant) {
  const int base = (ORACLE_PLANT_BASE0 == plant ? 0 : 1);
  const int kk = (ORACLE_PLANT_DROPK == plant ? k - 1 : k);
  const int browmajor = (ORACLE_PLANT_BLAYOUT == plant ? !b_rowmajor : b_rowmajor);
  long i;
  int p;
  if ((3 != type && 1 != type) || 0 > stack_size || 0 >= m || 0 >= n || 0 >= k) return 1;
  for (i = 0; i < c_elems; ++i) {
    c_hi[i] = (NULL != c_in ? load(c_in, type, i) : 0.0);
    c_lo[i] = 0.0;
  }
#if defined(_OPENMP)
#  pragma omp parallel for schedule(static)
#endif
  for (p = 0; p < m * n; ++p) { /* each thread owns one (row, column) position in every C block */
    const int r = p % m, col = p / m;
    int s;
    for (s = 0; s < stack_size; ++s) {
      const long a0 = (long)stack[3 * s + 0] - base, b0 = (long)stack[3 * s + 1] - base;
      const long c0 = (long)stack[3 * s + 2] - base;
      const long ci = c0 + (long)col * m + r;
      double hi = 0, lo = 0;
      int l;
      for (l = 0; l < kk; ++l) {
        const double x = load(a, type, a0 + (long)l * m + r);
        const double y = load(b, type, b0 + (0 != browmajor ? (long)l * n + col : (long)col * k + l));
        const double prod = x * y, err = fma(x, y, -prod);
        dd_add(&hi, &lo, prod, err);
      }
      if (0 > ci || ci >= c_elems) continue;
      if (ORACLE_PLANT_NOACC == plant) {
        c_hi[ci] = hi;
        c_lo[ci] = lo;
      }
      else {
        dd_add(c_hi + ci, c_lo + ci, hi, lo);
      }
    }
  }
  return 0;
}

/* out[0] = max|C-C_ref| / max(max|C_ref|, tiny), out[1] = ||C-C_ref||_F / ||C_ref||_F,
 * out[2] = max|C_ref|, out[3] = max|C-C_ref| */
int oracle_error(int type, long nelems, const void* test, const double* ref_hi, const double* ref_lo, double* out) {
  double maxref = 0, maxerr = 0, sumerr = 0, sumref = 0;
  long i;
  int nan = 0;
  for (i = 0; i < nelems; ++i) {
    const double t = load(test, type, i);
    const double d = (t - ref_hi[i]) - ref_lo[i];
    const double r = ref_hi[i] + ref_lo[i];
    if (fabs(r) > maxref) maxref = fabs(r);
    if (isnan(d) || isinf(d)) nan = 1; /* a non-finite result is an error of its own */
    else if (fabs(d) > maxerr) maxerr = fabs(d);
    sumerr += d * d;
    sumref += r * r;
  }
  if (0 != nan) maxerr = NAN;
  out[0] = maxerr / (maxref > 1e-300 ? maxref : 1e-300);
  out[1] = (sumref > 0 ? sqrt(sumerr / sumref) : sqrt(sumerr));
  out[2] = maxref;
  out[3] = maxerr;
  return 0;
}

int oracle_transpose(int type, int m, int n, const int* trs, int offset, int stack_size, void* data) {
  const size_t ts = (3 == type ? 8 : 4);
  unsigned char* const tmp = (unsigned char*)malloc(ts * (size_t)m * (size_t)n);
  int s, r, c;
  if (NULL == tmp) return 1;
  for (s = 0; s < stack_size; ++s) {
    unsigned char* const mat = (unsigned char*)data + ts * (size_t)trs[offset + s];
    memcpy(tmp, mat, ts * (size_t)m * (size_t)n);
    for (r = 0; r < m; ++r) {
      for (c = 0; c < n; ++c) memcpy(mat + ts * ((size_t)r * n + c), tmp + ts * ((size_t)c * m + r), ts);
    }
  }
  free(tmp);
  return 0;
}

int oracle_glibc_rand(unsigned int seed, int n, int* out) {
  int i;
  srand(seed);
  for (i = 0; i < n; ++i) out[i] = rand();
  return 0;
}

/* INIT_STACK (acc_bench.h:48-79) with a supplied rnd table (rnd_size > 0) */
int oracle_init_stack(int* stack, int stack_size, int rnd_size, const int* rnd, int mn, int mk, int kn, int nc, int na,
  int nb) {
  const int navg = stack_size / nc;
  const int nimb = (navg - 4 > 1 ? navg - 4 : 1);
  int i = 0, c = 0, ntop = 0;
  int* p = stack;
  if (0 >= nc || nc > stack_size || 0 >= rnd_size || NULL == rnd) return 1;
  while (i < stack_size) {
    const int r = rnd[i % rnd_size];
    const int next = c + 1;
    const int step = navg + (r % (2 * nimb) - nimb);
    ntop += (0 < step ? step : 1); /* the macro's step is 0 for an even r when navg = 1: it never ends there */
    if (stack_size < ntop) ntop = stack_size;
    for (; i < ntop; ++i) {
      const int ia = rnd[(2 * i + 0) % rnd_size] % na;
      const int ib = rnd[(2 * i + 1) % rnd_size] % nb;
      *p++ = ia * mk + 1;
      *p++ = ib * kn + 1;
      *p++ = c * mn + 1;
    }
    if (next < nc) c = next;
  }
  return 0;
}

/* INIT_MAT (acc_bench.h:30-40): column-major m-by-n, value (scale*seed + scale) * (index + 1) */
int oracle_init_mat(int type, double seed, void* mat, int m, int n, double scale) {
  const double seed1 = scale * seed + scale;
  int i, j;
  for (i = 0; i < n; ++i) {
    for (j = 0; j < m; ++j) {
      const int idx = i * m + j;
      if (3 == type) {
        ((double*)mat)[idx] = seed1 * (idx + 1);
      }
      else {
        ((float*)mat)[idx] = (float)(seed1 * (idx + 1));
      }
    }
  }
  return 0;
}

int oracle_max_threads(void) {
#if defined(_OPENMP)
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/*
 * GEMM oracle for the Ozaki emulation: C = alpha * op(A) * op(B) + beta * C, column-major BLAS
 * semantics (samples/ozaki/ozaki_opencl.h:420-421), every product exact (TwoProd) and summed in
 * double-double; alpha/beta applied in double-double too. Only the entries listed in (rows, cols)
 * are computed when nsample > 0 (size-independent spot checks of large products), else all.
 */
static inline void dd_mul_d(double hi, double lo, double b, double* rhi, double* rlo) {
  const double p = hi * b, e = fma(hi, b, -p);
  double s, t;
  s = p;
  t = e + lo * b;
  *rhi = s + t;
  *rlo = t - (*rhi - s);
}

int oracle_gemm_dd(char transa, char transb, int m, int n, int k, double alpha, const double* a, int lda, const double* b,
  int ldb, double beta, const double* c_in, int ldc, double* c_hi, double* c_lo, int nsample, const int* rows,
  const int* cols) {
  const long count = (0 < nsample ? nsample : (long)m * n);
  long x;
  if (0 > m || 0 > n || 0 > k) return 1;
#if defined(_OPENMP)
#  pragma omp parallel for schedule(dynamic, 16)
#endif
  for (x = 0; x < count; ++x) {
    const int i = (0 < nsample ? rows[x] : (int)(x % m)), j = (0 < nsample ? cols[x] : (int)(x / m));
    double hi = 0, lo = 0, rh, rl, ch, cl;
    int l;
    for (l = 0; l < k; ++l) {
      const double av = ('N' == transa || 'n' == transa) ? a[i + (long)l * lda] : a[l + (long)i * lda];
      const double bv = ('N' == transb || 'n' == transb) ? b[l + (long)j * ldb] : b[j + (long)l * ldb];
      const double prod = av * bv, err = fma(av, bv, -prod);
      dd_add(&hi, &lo, prod, err);
    }
    dd_mul_d(hi, lo, alpha, &rh, &rl);
    ch = 0;
    cl = 0;
    if (0 != beta && NULL != c_in) dd_mul_d(c_in[i + (long)j * ldc], 0.0, beta, &ch, &cl);
    dd_add(&rh, &rl, ch, cl);
    c_hi[x] = rh;
    c_lo[x] = rl;
  }
  return 0;
}
