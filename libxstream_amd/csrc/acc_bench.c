/*
 * acc_bench: the DBCSR/ACC stack benchmark of samples/smm/acc_bench.c, on this library's C interface.
 *
 *   acc_bench [nrepeat [stack_size [m [n [k [nc [na [nb]]]]]]]]
 *   acc_bench [nrepeat [stack_size [MxNxK [nc [na [nb]]]]]]
 *   acc_bench MxNxK [nc [na [nb]]]
 *   acc_bench <file>        one such argument list per line (blank lines skipped), one benchmark each
 *
 * Same arguments, defaults and workload as the reference (acc_bench.c:95-262, :491-557): m = 23, n = k = m,
 * stack_size 30000, nc = stack_size / 16, na = nb = 10 nc; the stack and the pools come from the reference's
 * generators (acc_bench.h:30-79) over a glibc rand() table seeded 25071975. stack_size with a k/m/g suffix
 * is a data budget: the smallest stack whose pools and stack take at least that many bytes; a negative
 * stack_size draws a random multiple of 100 up to its magnitude. For each shape it copies the pools and the
 * stack in from pinned host memory, transposes B on the device, runs libsmm_acc_process nrepeat times after
 * two warm-up calls, repeats the stacks on the host, and prints the reference's report lines (copy-in,
 * transpose, device, host, diff.cur; diff.max over a file's shapes). The host result is a multiply in
 * double (OpenMP threads own disjoint C blocks); diff.cur is the max normalised error
 * max |C - C_host| / max |C_host|.
 * Environment (acc_bench.c:96-117, :463-477): CHECK=<factor> fails a shape when diff.cur exceeds factor x
 * 1e-3 (FP64) or 2e-3 (FP32); CHECK=0 skips the host run, repeats 66 times by default and lets a file
 * carry on past shapes without a kernel. CHECK_H2D, CHECK_DEV, CHECK_HST: floors (GB/s, GFLOPS/s) that
 * the last shape's copy-in, device and host rates must reach. NREPEAT_H2D (copy-in repetitions),
 * BATCHSIZE_SMM=<stack_size>[,<nrepeat>] (overrides the argument), DEVICE (device id).
 * Built twice: double (acc_bench) and float (acc_bench_f32, -DELEM_TYPE=float).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#if defined(_OPENMP)
#  include <omp.h>
#endif
#include "../../include/libxstream_amd.h"

#if !defined(ELEM_TYPE)
#  define ELEM_TYPE double
#endif
#define BENCH_GRAIN 100                    /* granularity of a random stack size */
#define BENCH_BATCHSIZE (300 * BENCH_GRAIN) /* default stack size */
#define BENCH_NRAND BENCH_BATCHSIZE
#define BENCH_NREPEAT 3   /* repetitions with a host check */
#define BENCH_XREPEAT 66  /* repetitions without (CHECK=0) */
#define BENCH_WARMUP 2
#define BENCH_MAX_KERNEL_DIM 80
#define BENCH_DELIMS ";,:|/\t\r\n "

static double bench_now(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* a column-major rows x cols block: element i = scale (seed + 1) (i + 1) (acc_bench.h:30-40) */
static void bench_init_mat(ELEM_TYPE* mat, int seed, int rows, int cols, double scale) {
  const double s1 = scale * seed + scale;
  int i;
  for (i = 0; i < rows * cols; ++i) mat[i] = (ELEM_TYPE)(s1 * (i + 1));
}

/* runs of stack_size / nc +- (that - 4) entries per C block, a and b drawn from the table (acc_bench.h:48-79);
   a run takes at least one entry: with stack_size / nc = 1 the macro's step can be 0 and it never ends */
static void bench_init_stack(int* stack, int stack_size, const int* rnd, int nrnd, int mn, int mk, int kn, int nc,
  int na, int nb) {
  const int navg = stack_size / nc, nimb = (1 < navg - 4 ? navg - 4 : 1);
  int i = 0, c = 0, ntop = 0;
  while (i < stack_size) {
    const int r = rnd[i % nrnd], next = c + 1;
    const int step = navg + (r % (2 * nimb) - nimb);
    ntop += (0 < step ? step : 1);
    if (stack_size < ntop) ntop = stack_size;
    for (; i < ntop; ++i) {
      stack[3 * i + 0] = (rnd[(2 * i + 0) % nrnd] % na) * mk + 1;
      stack[3 * i + 1] = (rnd[(2 * i + 1) % nrnd] % nb) * kn + 1;
      stack[3 * i + 2] = c * mn + 1;
    }
    if (next < nc) c = next;
  }
}

/* one argument list: the strings as given (NULL: absent) */
typedef struct bench_args_t {
  const char *nrepeat, *stack_size, *m, *n, *k, *nc, *na, *nb;
} bench_args_t;

/* split an argument list into its fields; "MxNxK" stands for m, n and k (acc_bench.c:491-557) */
static void bench_fields(int argc, char* args[], bench_args_t* a) {
  int i = 0;
  memset(a, 0, sizeof(*a));
  if (0 < argc) {
    const char* const x1 = strchr(args[0], 'x');
    if (NULL == x1 || NULL == strchr(x1 + 1, 'x')) {
      a->nrepeat = args[i++];
      if (i < argc) a->stack_size = args[i++];
    }
  }
  if (i < argc) {
    const char* const x1 = strchr(args[i], 'x');
    const char* const x2 = (NULL != x1 ? strchr(x1 + 1, 'x') : NULL);
    a->m = args[i++];
    if (NULL != x2) {
      a->n = x1 + 1;
      a->k = x2 + 1;
    }
    else {
      if (i < argc) a->n = args[i++];
      if (i < argc) a->k = args[i++];
    }
  }
  if (i < argc) a->nc = args[i++];
  if (i < argc) a->na = args[i++];
  if (i < argc) a->nb = args[i++];
}

/* the next non-blank line of the file as an argument list; 0 at the end (or on a line of more than 8 fields) */
static int bench_next_line(FILE* file, char* line, size_t size, bench_args_t* a) {
  while (NULL != fgets(line, (int)size, file)) {
    char *args[8], *save = NULL, *tok;
    int argc = 0;
    for (tok = strtok_r(line, BENCH_DELIMS, &save); NULL != tok; tok = strtok_r(NULL, BENCH_DELIMS, &save)) {
      if (8 <= argc) return 0;
      args[argc++] = tok;
    }
    if (0 < argc) {
      bench_fields(argc, args, a);
      return 1;
    }
  }
  return 0;
}

static int bench_int(const char* s) {
  return (NULL != s ? atoi(s) : 0);
}

/* sizes that follow from the stack size and the given nc, na, nb (0: default) */
static void bench_pools(int stack_size, int inc, int ina, int inb, int* nc, int* na, int* nb) {
  *nc = (0 < inc ? (inc < stack_size ? inc : stack_size) : (1 < stack_size / 16 ? stack_size / 16 : 1));
  *na = (0 < ina ? ina : 10 * *nc);
  *nb = (0 < inb ? inb : 10 * *nc);
}

/* stack size from its argument: plain count, a byte budget (k/m/g suffix), or <= 0 (default/random) */
static int bench_stack_size(const char* arg, int mk, int kn, int inc, int ina, int inb) {
  const int ts = (int)sizeof(ELEM_TYPE);
  char* end = NULL;
  const long v = (NULL != arg ? strtol(arg, &end, 10) : 0);
  const char* const units = "kmgKMG";
  const char* const unit = (NULL != end && '\0' != *end && '\0' == end[1]) ? strchr(units, *end) : NULL;
  if (NULL != unit && 0 < v) { /* smallest stack whose pools and stack take >= budget bytes */
    const double budget = (double)v * (double)(1ULL << (10 * ((unit - units) % 3 + 1)));
    long lo = 0, hi = 1;
    for (;;) { /* the data size grows with the stack size: bracket, then bisect */
      int nc, na, nb;
      bench_pools((int)hi, inc, ina, inb, &nc, &na, &nb);
      if (budget <= (double)ts * ((double)mk * na + (double)kn * nb) + 12.0 * hi || (1L << 30) <= hi) break;
      lo = hi;
      hi *= 2;
    }
    while (lo < hi) {
      const long mid = lo + (hi - lo) / 2;
      int nc, na, nb;
      bench_pools((int)mid, inc, ina, inb, &nc, &na, &nb);
      if (budget <= (double)ts * ((double)mk * na + (double)kn * nb) + 12.0 * mid) hi = mid;
      else lo = mid + 1;
    }
    return (int)lo;
  }
  return (NULL != end && '\0' == *end) ? (int)v : 0;
}

typedef struct bench_perf_t {
  double h2d, dev, hst, diff; /* GB/s, GFLOPS/s, GFLOPS/s, diff.cur (-1: none) */
} bench_perf_t;

#define BENCH_CHECK(EXPR, WHAT) \
  do { \
    if (EXIT_SUCCESS == result) { \
      result = (EXPR); \
      if (EXIT_SUCCESS != result && !(0 > result && 0 == check)) fprintf(stderr, "ERROR: %s failed (%i)\n", (WHAT), result); \
    } \
  } while (0)

/* report lines of one benchmark, printed only when it succeeds (acc_bench.c:447-452) */
static char bench_report[4096];
static size_t bench_nreport;
#define BENCH_PRINTF(...) \
  do { \
    if (bench_nreport < sizeof(bench_report)) { \
      const int bench_n_ = snprintf(bench_report + bench_nreport, sizeof(bench_report) - bench_nreport, __VA_ARGS__); \
      if (0 < bench_n_) bench_nreport += (size_t)bench_n_; \
    } \
  } while (0)

/* one benchmark: 0 on success, -1 when no kernel serves the shape, else EXIT_FAILURE */
static int bench_one(const char* name, const bench_args_t* a, const int* rnd, int nok, double check, bench_perf_t* perf) {
  const int ts = (int)sizeof(ELEM_TYPE), tid = (8 == ts ? dbcsr_type_real_8 : dbcsr_type_real_4);
  const char *const env_h2d = getenv("NREPEAT_H2D"), *const env_batch = getenv("BATCHSIZE_SMM");
  const int nrepeat_h2d = (NULL != env_h2d && 0 < atoi(env_h2d) ? atoi(env_h2d) : 1);
  const int xrepeat = (0 != check ? BENCH_NREPEAT : BENCH_XREPEAT);
  const int m = (0 < bench_int(a->m) ? bench_int(a->m) : 23);
  const int n = (0 < bench_int(a->n) ? bench_int(a->n) : m), k = (0 < bench_int(a->k) ? bench_int(a->k) : m);
  const int inc = bench_int(a->nc), ina = bench_int(a->na), inb = bench_int(a->nb);
  const int mn = m * n, mk = m * k, kn = k * n;
  int nrepeat = (0 < bench_int(a->nrepeat) ? bench_int(a->nrepeat) : xrepeat), stack_size, nc, na, nb, nr, i, r;
  int *stack_hst = NULL, *trans_hst = NULL, *stack_dev = NULL, *trans_dev = NULL;
  ELEM_TYPE *a_hst = NULL, *b_hst = NULL, *c_hst = NULL, *a_dev = NULL, *b_dev = NULL, *c_dev = NULL;
  double t0, transpose = 0, duration = 0;
  void* stream = NULL;
  int result = EXIT_SUCCESS;
  perf->h2d = perf->dev = perf->hst = 0;
  perf->diff = -1;
  if (NULL == env_batch) {
    stack_size = bench_stack_size(a->stack_size, mk, kn, inc, ina, inb);
  }
  else { /* BATCHSIZE_SMM=<stack_size>[<delimiter><nrepeat>] */
    const size_t sep = strcspn(env_batch, BENCH_DELIMS);
    if ('\0' != env_batch[sep] && 0 < atoi(env_batch + sep + 1)) nrepeat = atoi(env_batch + sep + 1);
    stack_size = atoi(env_batch);
  }
  if (0 > stack_size) { /* random multiple of BENCH_GRAIN up to |stack_size|, same total work as the default */
    const int ss = -stack_size, bs = (1 < ss ? ss : BENCH_BATCHSIZE);
    const int limit = (BENCH_GRAIN < ss ? (bs + BENCH_GRAIN - 1) / BENCH_GRAIN : ss);
    const long work = (long)BENCH_BATCHSIZE * nrepeat;
    stack_size = (rnd[nok % BENCH_NRAND] % limit + 1) * BENCH_GRAIN;
    nrepeat = (int)((work + stack_size - 1) / stack_size);
    if (nrepeat < xrepeat) nrepeat = xrepeat;
  }
  else if (0 == stack_size) {
    stack_size = BENCH_BATCHSIZE;
  }
  bench_pools(stack_size, inc, ina, inb, &nc, &na, &nb);
  nr = nrepeat * nc;
  BENCH_PRINTF("%s %i %i %i %i %i %i %i %i\n", name, nrepeat, stack_size, m, n, k, nc, na, nb);
  BENCH_PRINTF("typename (id=%i): %s\n", tid, 8 == ts ? "double" : "float");
  if (BENCH_MAX_KERNEL_DIM < (int)ceil(sqrt((double)mn))) {
    fprintf(stderr, "ERROR: Matrix shape (%i) exceeds MAX_KERNEL_DIM=%i\n", (int)ceil(sqrt((double)mn)),
      BENCH_MAX_KERNEL_DIM);
    return EXIT_FAILURE;
  }
  BENCH_CHECK(c_dbcsr_acc_stream_create(&stream, "stream", -1), "stream create");
  BENCH_CHECK(c_dbcsr_acc_host_mem_allocate((void**)&a_hst, (size_t)ts * mk * na, stream), "host alloc");
  BENCH_CHECK(c_dbcsr_acc_host_mem_allocate((void**)&b_hst, (size_t)ts * kn * nb, stream), "host alloc");
  BENCH_CHECK(c_dbcsr_acc_host_mem_allocate((void**)&c_hst, (size_t)ts * mn * nc, stream), "host alloc");
  BENCH_CHECK(c_dbcsr_acc_host_mem_allocate((void**)&stack_hst, sizeof(int) * 3 * (size_t)stack_size, stream), "host alloc");
  BENCH_CHECK(c_dbcsr_acc_host_mem_allocate((void**)&trans_hst, sizeof(int) * (size_t)nb, stream), "host alloc");
  BENCH_CHECK(c_dbcsr_acc_stream_sync(stream), "stream sync");
  if (EXIT_SUCCESS == result) {
    bench_init_stack(stack_hst, stack_size, rnd, BENCH_NRAND, mn, mk, kn, nc, na, nb);
    for (i = 0; i < na; ++i) bench_init_mat(a_hst + (size_t)i * mk, i + 42, m, k, 1.0 / ((double)nr * na));
    for (i = 0; i < nb; ++i) {
      bench_init_mat(b_hst + (size_t)i * kn, i + 24, k, n, 1.0 / ((double)nr * nb));
      trans_hst[i] = i * kn;
    }
  }
  BENCH_CHECK(c_dbcsr_acc_dev_mem_allocate((void**)&a_dev, (size_t)ts * mk * na), "device alloc");
  BENCH_CHECK(c_dbcsr_acc_dev_mem_allocate((void**)&b_dev, (size_t)ts * kn * nb), "device alloc");
  BENCH_CHECK(c_dbcsr_acc_dev_mem_allocate((void**)&c_dev, (size_t)ts * mn * nc), "device alloc");
  BENCH_CHECK(c_dbcsr_acc_dev_mem_allocate((void**)&stack_dev, sizeof(int) * 3 * (size_t)stack_size), "device alloc");
  BENCH_CHECK(c_dbcsr_acc_dev_mem_allocate((void**)&trans_dev, sizeof(int) * (size_t)nb), "device alloc");
  BENCH_CHECK(c_dbcsr_acc_memset_zero(c_dev, 0, (size_t)ts * mn * nc, stream), "memset");
  BENCH_CHECK(c_dbcsr_acc_memcpy_h2d(trans_hst, trans_dev, sizeof(int) * (size_t)nb, stream), "copy-in");
  BENCH_CHECK(c_dbcsr_acc_stream_sync(stream), "stream sync");
  t0 = bench_now();
  for (r = 0; r < nrepeat_h2d && EXIT_SUCCESS == result; ++r) { /* acc_bench.c:288-307 */
    BENCH_CHECK(c_dbcsr_acc_memcpy_h2d(a_hst, a_dev, (size_t)ts * mk * na, stream), "copy-in");
    BENCH_CHECK(c_dbcsr_acc_memcpy_h2d(b_hst, b_dev, (size_t)ts * kn * nb, stream), "copy-in");
    BENCH_CHECK(c_dbcsr_acc_memcpy_h2d(stack_hst, stack_dev, sizeof(int) * 3 * (size_t)stack_size, stream), "copy-in");
  }
  BENCH_CHECK(c_dbcsr_acc_stream_sync(stream), "stream sync");
  if (EXIT_SUCCESS == result) {
    const double size = (double)ts * ((double)mk * na + (double)kn * nb) + 12.0 * stack_size;
    const double dt = (bench_now() - t0) / nrepeat_h2d;
    perf->h2d = size / (dt * (1 << 30));
    BENCH_PRINTF("copy-in (%i MB): %.2g ms %.1f GB/s\n", (int)((size + (1 << 19)) / (1 << 20)), 1e3 * dt, perf->h2d);
  }
  /* B into the kernel layout: warm-up pairs, then one timed transpose (acc_bench.c:313-327) */
  for (r = 0; r < BENCH_WARMUP / 2; ++r) {
    BENCH_CHECK(libsmm_acc_transpose(trans_dev, 0, nb, b_dev, (libsmm_acc_data_t)tid, k, n, BENCH_MAX_KERNEL_DIM, stream),
      "transpose");
    BENCH_CHECK(libsmm_acc_transpose(trans_dev, 0, nb, b_dev, (libsmm_acc_data_t)tid, n, k, BENCH_MAX_KERNEL_DIM, stream),
      "transpose");
  }
  BENCH_CHECK(c_dbcsr_acc_stream_sync(stream), "stream sync");
  t0 = bench_now();
  BENCH_CHECK(libsmm_acc_transpose(trans_dev, 0, nb, b_dev, (libsmm_acc_data_t)tid, k, n, BENCH_MAX_KERNEL_DIM, stream),
    "transpose");
  BENCH_CHECK(c_dbcsr_acc_stream_sync(stream), "stream sync");
  transpose = bench_now() - t0;
  /* warm-up, then nrepeat timed applications of the stack to zeroed C (acc_bench.c:328-346) */
  for (r = 0; r < BENCH_WARMUP; ++r) {
    BENCH_CHECK(libsmm_acc_process(stack_hst, stack_dev, stack_size, (libsmm_acc_data_t)tid, a_dev, b_dev, c_dev, m, n,
                  k, BENCH_MAX_KERNEL_DIM, 1, stream, stream), "libsmm_acc_process");
  }
  BENCH_CHECK(c_dbcsr_acc_memset_zero(c_dev, 0, (size_t)ts * mn * nc, stream), "memset");
  BENCH_CHECK(c_dbcsr_acc_stream_sync(stream), "stream sync");
  t0 = bench_now();
  for (r = 0; r < nrepeat && EXIT_SUCCESS == result; ++r) {
    BENCH_CHECK(libsmm_acc_process(stack_hst, stack_dev, stack_size, (libsmm_acc_data_t)tid, a_dev, b_dev, c_dev, m, n,
                  k, BENCH_MAX_KERNEL_DIM, 1, stream, stream), "libsmm_acc_process");
  }
  BENCH_CHECK(c_dbcsr_acc_stream_sync(stream), "stream sync");
  duration = bench_now() - t0;
  if (EXIT_SUCCESS == result) {
    const double gflop = 2e-9 * m * n * k * (double)stack_size * nrepeat;
    if (0 < duration) {
      perf->dev = gflop / duration;
      BENCH_PRINTF("transpose: %.2g ms %.1f GFLOPS/s\n", 1e3 * (duration + transpose) / nrepeat,
        gflop / (duration + transpose));
      BENCH_PRINTF("device: %.2g ms %.1f GFLOPS/s\n", 1e3 * duration / nrepeat, perf->dev);
    }
    else {
      BENCH_PRINTF("transpose: 0 ms 0 GFLOPS/s\n");
      BENCH_PRINTF("device: 0 ms 0 GFLOPS/s\n");
    }
  }
  /* the same nrepeat applications on the host (skipped for CHECK=0): B untransposed (column-major k x n), in
     double; every thread owns the C blocks c with c mod nthreads = its id, so any stack order is race-free */
  if (EXIT_SUCCESS == result && 0 != check) {
    double* const gold = (double*)malloc(sizeof(double) * (size_t)mn * nc);
    const double gflop = 2e-9 * m * n * k * (double)stack_size * nrepeat;
    if (NULL != gold) {
      double dt = 0, cmax = 0, dmax = 0, vref = 0, vtst = 0;
      int pass;
      for (pass = 0; pass < 2; ++pass) { /* warm-up pass (BENCH_WARMUP applications), then the timed one */
        const int napply = (0 == pass ? BENCH_WARMUP : nrepeat);
        memset(gold, 0, sizeof(double) * (size_t)mn * nc);
        t0 = bench_now();
#if defined(_OPENMP)
#  pragma omp parallel private(r)
#endif
        {
#if defined(_OPENMP)
          const int tid_omp = omp_get_thread_num(), nthreads = omp_get_num_threads();
#else
          const int tid_omp = 0, nthreads = 1;
#endif
          int x;
          for (r = 0; r < napply; ++r) {
            for (x = 0; x < stack_size; ++x) {
              const int cofs = stack_hst[3 * x + 2] - 1;
              if ((cofs / mn) % nthreads == tid_omp) {
                const ELEM_TYPE* const ab = a_hst + stack_hst[3 * x] - 1;
                const ELEM_TYPE* const bb = b_hst + stack_hst[3 * x + 1] - 1;
                double* const cb = gold + cofs;
                int ii, jj, kk;
                for (jj = 0; jj < n; ++jj) {
                  for (kk = 0; kk < k; ++kk) {
                    const double bv = (double)bb[kk + jj * k];
                    for (ii = 0; ii < m; ++ii) cb[ii + jj * m] += (double)ab[ii + kk * m] * bv;
                  }
                }
              }
            }
          }
        }
        dt = bench_now() - t0;
      }
      perf->hst = gflop / dt;
      BENCH_PRINTF("host: %.2g ms %.1f GFLOPS/s\n", 1e3 * dt / nrepeat, perf->hst);
      BENCH_CHECK(c_dbcsr_acc_memcpy_d2h(c_dev, c_hst, (size_t)ts * mn * nc, stream), "copy-out");
      BENCH_CHECK(c_dbcsr_acc_stream_sync(stream), "stream sync");
      if (EXIT_SUCCESS == result) {
        size_t e;
        for (e = 0; e < (size_t)mn * nc; ++e) {
          const double d = fabs((double)c_hst[e] - gold[e]);
          if (cmax < fabs(gold[e])) cmax = fabs(gold[e]);
          if (dmax < d) {
            dmax = d;
            vref = gold[e];
            vtst = (double)c_hst[e];
          }
        }
        perf->diff = (0 < cmax ? dmax / cmax : dmax);
        if (0 < perf->diff) BENCH_PRINTF("diff.cur: %g (|%g-%g|=%g)\n", perf->diff, vref, vtst, dmax);
        else BENCH_PRINTF("diff.cur: %g\n", perf->diff);
        if (0 < check && check < perf->diff) result = EXIT_FAILURE;
      }
      free(gold);
    }
    else {
      result = EXIT_FAILURE;
    }
  }
  c_dbcsr_acc_host_mem_deallocate(stack_hst, stream);
  c_dbcsr_acc_host_mem_deallocate(trans_hst, stream);
  c_dbcsr_acc_host_mem_deallocate(a_hst, stream);
  c_dbcsr_acc_host_mem_deallocate(b_hst, stream);
  c_dbcsr_acc_host_mem_deallocate(c_hst, stream);
  c_dbcsr_acc_dev_mem_deallocate(stack_dev);
  c_dbcsr_acc_dev_mem_deallocate(trans_dev);
  c_dbcsr_acc_dev_mem_deallocate(a_dev);
  c_dbcsr_acc_dev_mem_deallocate(b_dev);
  c_dbcsr_acc_dev_mem_deallocate(c_dev);
  if (NULL != stream) c_dbcsr_acc_stream_destroy(stream);
  if (0 > result) fprintf(stderr, "WARNING: %ix%ix%i-kernel not available or unsuitable!\n", m, n, k);
  return result;
}

int main(int argc, char* argv[]) {
  const char* const env_check = getenv("CHECK");
  const char* const env_device = getenv("DEVICE");
  const char *const env_check_h2d = getenv("CHECK_H2D"), *const env_check_dev = getenv("CHECK_DEV");
  const char* const env_check_hst = getenv("CHECK_HST");
  const int ts = (int)sizeof(ELEM_TYPE);
  const double check = (NULL != env_check ? fabs(atof(env_check)) * (8 == ts ? 1e-3 : 2e-3) : -1);
  FILE* file = (1 < argc ? fopen(argv[1], "r") : NULL);
  char line[1024];
  bench_args_t args;
  bench_perf_t perf = {0, 0, 0, -1};
  double maxdiff = 0;
  int *rnd = NULL, ndevices = 0, nok = 0, i;
  int result = c_dbcsr_acc_init();
  if (EXIT_SUCCESS == result) result = libsmm_acc_init();
  if (EXIT_SUCCESS == result) result = c_dbcsr_acc_get_ndevices(&ndevices);
  if (EXIT_SUCCESS != result || 0 >= ndevices) {
    fprintf(stderr, "ERROR: No ACC-device found!\n");
    if (NULL != file) fclose(file);
    return EXIT_FAILURE;
  }
  {
    const int device = (NULL != env_device && 0 <= atoi(env_device) && atoi(env_device) < ndevices ? atoi(env_device) : 0);
    result = c_dbcsr_acc_set_active_device(device);
    if (EXIT_SUCCESS == result) printf("Activated device%i (ndevices=%i)\n", device, ndevices);
    else fprintf(stderr, "ERROR: Failed to activate device!\n");
  }
  rnd = (int*)malloc(sizeof(int) * BENCH_NRAND);
  if (NULL == rnd) result = EXIT_FAILURE;
  if (EXIT_SUCCESS == result) {
    srand(25071975);
    for (i = 0; i < BENCH_NRAND; ++i) rnd[i] = rand();
    if (NULL != file) {
      if (0 == bench_next_line(file, line, sizeof(line), &args)) {
        fclose(file);
        file = NULL;
        bench_fields(0, NULL, &args);
      }
    }
    else {
      bench_fields(argc - 1, argv + 1, &args);
    }
  }
  while (EXIT_SUCCESS == result || (NULL != file && 0 > result && 0 == check)) {
    bench_nreport = 0;
    bench_report[0] = '\0';
    result = bench_one(argv[0], &args, rnd, nok, check, &perf);
    if (0 == result || (0 > result && 0 == check)) {
      if (0 <= perf.diff && maxdiff < perf.diff && NULL != file) maxdiff = perf.diff;
      if (NULL != file) {
        BENCH_PRINTF("\n");
        if (0 == bench_next_line(file, line, sizeof(line), &args)) {
          fclose(file);
          file = NULL;
        }
      }
      ++nok;
    }
    if (0 == result) fputs(bench_report, stdout);
    if (NULL == file) break;
  }
  if (NULL != file) fclose(file);
  free(rnd);
  libsmm_acc_finalize();
  c_dbcsr_acc_finalize();
  if (1 < nok) printf("\ndiff.max: %g\n", maxdiff);
  if (EXIT_SUCCESS == result) { /* performance floors of the last benchmark (acc_bench.c:463-477) */
    if ((NULL != env_check_h2d && 0 < perf.h2d && perf.h2d < atof(env_check_h2d)) ||
        (NULL != env_check_dev && 0 < perf.dev && perf.dev < atof(env_check_dev)) ||
        (NULL != env_check_hst && 0 < perf.hst && perf.hst < atof(env_check_hst))) {
      result = EXIT_FAILURE;
    }
    if (EXIT_SUCCESS != result) fprintf(stderr, "\nFAILED\n\n");
  }
  else if (0 < result) {
    fprintf(stderr, "\nFAILED\n\n");
  }
  else if (0 == check) {
    result = EXIT_SUCCESS;
  }
  return (EXIT_SUCCESS == result ? EXIT_SUCCESS : EXIT_FAILURE);
}
