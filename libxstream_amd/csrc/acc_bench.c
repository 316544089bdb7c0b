/*
 * acc_bench: the DBCSR/ACC stack benchmark of samples/smm/acc_bench.c, on this library's C interface.
 *
 *   acc_bench [nrepeat [stack_size [m [n [k [nc [na [nb]]]]]]]]
 *
 * Same arguments, defaults and workload as the reference (acc_bench.c:165-262): m = 23, n = k = m,
 * stack_size 30000, nc = stack_size / 16, na = nb = 10 nc, 3 repetitions; the stack and the pools come from
 * the reference's generators (acc_bench.h:30-79) over a glibc rand() table seeded 25071975. It copies the
 * pools and the stack in from pinned host memory, transposes B on the device, runs libsmm_acc_process
 * nrepeat times, repeats the stacks on the host, and prints the same report lines (copy-in, transpose,
 * device, host, diff.cur). The host result is a plain multiply in double; diff.cur is the max normalised
 * error max |C - C_host| / max |C_host|. Environment: CHECK=<factor> fails the run when diff.cur exceeds
 * factor x 1e-3 (FP64) or 2e-3 (FP32) (acc_bench.c:65-67, :96-97), NREPEAT_H2D (copy-in repetitions),
 * DEVICE (device id). Built twice: double (acc_bench) and float (acc_bench_f32, -DELEM_TYPE=float).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "../../include/libxstream_amd.h"

#if !defined(ELEM_TYPE)
#  define ELEM_TYPE double
#endif
#define BENCH_BATCHSIZE 30000
#define BENCH_NRAND BENCH_BATCHSIZE
#define BENCH_NREPEAT 3
#define BENCH_MAX_KERNEL_DIM 80

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

#define BENCH_CHECK(EXPR, WHAT) \
  do { \
    if (EXIT_SUCCESS == result) { \
      result = (EXPR); \
      if (EXIT_SUCCESS != result && -1 != result) fprintf(stderr, "ERROR: %s failed (%i)\n", (WHAT), result); \
    } \
  } while (0)

int main(int argc, char* argv[]) {
  const int ts = (int)sizeof(ELEM_TYPE), tid = (8 == ts ? dbcsr_type_real_8 : dbcsr_type_real_4);
  const char *const env_check = getenv("CHECK"), *const env_h2d = getenv("NREPEAT_H2D");
  const char* const env_device = getenv("DEVICE");
  const double epsilon = (8 == ts ? 1e-3 : 2e-3);
  const double check = (NULL != env_check ? fabs(atof(env_check)) * epsilon : -1);
  const int nrepeat_h2d = (NULL != env_h2d && 0 < atoi(env_h2d) ? atoi(env_h2d) : 1);
  const int nrepeat = (1 < argc && 0 < atoi(argv[1]) ? atoi(argv[1]) : BENCH_NREPEAT);
  const int stack_size = (2 < argc && 0 < atoi(argv[2]) ? atoi(argv[2]) : BENCH_BATCHSIZE);
  const int m = (3 < argc && 0 < atoi(argv[3]) ? atoi(argv[3]) : 23);
  const int n = (4 < argc && 0 < atoi(argv[4]) ? atoi(argv[4]) : m);
  const int k = (5 < argc && 0 < atoi(argv[5]) ? atoi(argv[5]) : m);
  const int nc = (6 < argc && 0 < atoi(argv[6]) ? (atoi(argv[6]) < stack_size ? atoi(argv[6]) : stack_size)
                                                 : (1 < stack_size / 16 ? stack_size / 16 : 1));
  const int na = (7 < argc && 0 < atoi(argv[7]) ? atoi(argv[7]) : 10 * nc);
  const int nb = (8 < argc && 0 < atoi(argv[8]) ? atoi(argv[8]) : 10 * nc);
  const int mn = m * n, mk = m * k, kn = k * n, nr = nrepeat * nc;
  int *rnd = NULL, *stack_hst = NULL, *trans_hst = NULL, *stack_dev = NULL, *trans_dev = NULL;
  ELEM_TYPE *a_hst = NULL, *b_hst = NULL, *c_hst = NULL, *a_dev = NULL, *b_dev = NULL, *c_dev = NULL;
  double *gold = NULL, t0, transpose = 0, duration = 0, diff = 0;
  void* stream = NULL;
  int result = c_dbcsr_acc_init(), ndevices = 0, i, r;
  BENCH_CHECK(libsmm_acc_init(), "libsmm_acc_init");
  BENCH_CHECK(c_dbcsr_acc_get_ndevices(&ndevices), "c_dbcsr_acc_get_ndevices");
  if (EXIT_SUCCESS != result || 0 >= ndevices) {
    fprintf(stderr, "ERROR: No ACC-device found!\n");
    return EXIT_FAILURE;
  }
  {
    const int device = (NULL != env_device && 0 <= atoi(env_device) && atoi(env_device) < ndevices ? atoi(env_device) : 0);
    BENCH_CHECK(c_dbcsr_acc_set_active_device(device), "c_dbcsr_acc_set_active_device");
    if (EXIT_SUCCESS == result) printf("Activated device%i (ndevices=%i)\n", device, ndevices);
  }
  printf("%s %i %i %i %i %i %i %i %i\n", argv[0], nrepeat, stack_size, m, n, k, nc, na, nb);
  printf("typename (id=%i): %s\n", tid, 8 == ts ? "double" : "float");
  if (BENCH_MAX_KERNEL_DIM < (int)ceil(sqrt((double)mn))) {
    fprintf(stderr, "ERROR: Matrix shape (%i) exceeds MAX_KERNEL_DIM=%i\n", (int)ceil(sqrt((double)mn)),
      BENCH_MAX_KERNEL_DIM);
    return EXIT_FAILURE;
  }
  rnd = (int*)malloc(sizeof(int) * BENCH_NRAND);
  if (NULL == rnd) return EXIT_FAILURE;
  srand(25071975);
  for (i = 0; i < BENCH_NRAND; ++i) rnd[i] = rand();
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
    printf("copy-in (%i MB): %.2g ms %.1f GB/s\n", (int)((size + (1 << 19)) / (1 << 20)), 1e3 * dt,
      size / (dt * (1 << 30)));
  }
  /* B into the kernel layout: warm-up pair, then one timed transpose (acc_bench.c:313-327) */
  BENCH_CHECK(libsmm_acc_transpose(trans_dev, 0, nb, b_dev, (libsmm_acc_data_t)tid, k, n, BENCH_MAX_KERNEL_DIM, stream),
    "transpose");
  BENCH_CHECK(libsmm_acc_transpose(trans_dev, 0, nb, b_dev, (libsmm_acc_data_t)tid, n, k, BENCH_MAX_KERNEL_DIM, stream),
    "transpose");
  BENCH_CHECK(c_dbcsr_acc_stream_sync(stream), "stream sync");
  t0 = bench_now();
  BENCH_CHECK(libsmm_acc_transpose(trans_dev, 0, nb, b_dev, (libsmm_acc_data_t)tid, k, n, BENCH_MAX_KERNEL_DIM, stream),
    "transpose");
  BENCH_CHECK(c_dbcsr_acc_stream_sync(stream), "stream sync");
  transpose = bench_now() - t0;
  /* warm-up, then nrepeat timed applications of the stack to zeroed C (acc_bench.c:328-346) */
  BENCH_CHECK(libsmm_acc_process(stack_hst, stack_dev, stack_size, (libsmm_acc_data_t)tid, a_dev, b_dev, c_dev, m, n, k,
                BENCH_MAX_KERNEL_DIM, 1, stream, stream), "libsmm_acc_process");
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
    printf("transpose: %.2g ms %.1f GFLOPS/s\n", 1e3 * (duration + transpose) / nrepeat, gflop / (duration + transpose));
    printf("device: %.2g ms %.1f GFLOPS/s\n", 1e3 * duration / nrepeat, gflop / duration);
  }
  /* the same nrepeat applications on the host: B untransposed (column-major k x n), in double */
  gold = (double*)calloc((size_t)mn * nc, sizeof(double));
  if (EXIT_SUCCESS == result && NULL != gold) {
    const double gflop = 2e-9 * m * n * k * (double)stack_size * nrepeat;
    double dt, cmax = 0;
    int x;
    t0 = bench_now();
    for (r = 0; r < nrepeat; ++r) {
      for (x = 0; x < stack_size; ++x) {
        const ELEM_TYPE* const ab = a_hst + stack_hst[3 * x] - 1;
        const ELEM_TYPE* const bb = b_hst + stack_hst[3 * x + 1] - 1;
        double* const cb = gold + stack_hst[3 * x + 2] - 1;
        int ii, jj, kk;
        for (jj = 0; jj < n; ++jj) {
          for (kk = 0; kk < k; ++kk) {
            const double bv = (double)bb[kk + jj * k];
            for (ii = 0; ii < m; ++ii) cb[ii + jj * m] += (double)ab[ii + kk * m] * bv;
          }
        }
      }
    }
    dt = bench_now() - t0;
    printf("host: %.2g ms %.1f GFLOPS/s\n", 1e3 * dt / nrepeat, gflop / dt);
    BENCH_CHECK(c_dbcsr_acc_memcpy_d2h(c_dev, c_hst, (size_t)ts * mn * nc, stream), "copy-out");
    BENCH_CHECK(c_dbcsr_acc_stream_sync(stream), "stream sync");
    if (EXIT_SUCCESS == result) {
      size_t e;
      for (e = 0; e < (size_t)mn * nc; ++e) {
        const double d = fabs((double)c_hst[e] - gold[e]);
        if (cmax < fabs(gold[e])) cmax = fabs(gold[e]);
        if (diff < d) diff = d;
      }
      diff = (0 < cmax ? diff / cmax : diff);
      printf("diff.cur: %g\n", diff);
      if (0 < check && check < diff) result = EXIT_FAILURE;
    }
  }
  free(gold);
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
  free(rnd);
  libsmm_acc_finalize();
  c_dbcsr_acc_finalize();
  if (-1 == result) fprintf(stderr, "WARNING: %ix%ix%i-kernel not available or unsuitable!\n", m, n, k);
  return (EXIT_SUCCESS == result ? EXIT_SUCCESS : EXIT_FAILURE);
}
