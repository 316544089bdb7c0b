/*
 * Host-layer concurrency under ThreadSanitizer (no GPU needed): the configuration registry, the
 * learned-predictor slot and the thread-local error state, driven from 8 threads at once the way
 * concurrent libsmm_acc_process calls drive them (samples/smm/smm_kernel.c:91,109-113: the reference
 * guards its registry with lock stripes; libsmm_acc_is_thread_safe returns 1). Built with
 * -fsanitize=thread, and again with -fsanitize=address,undefined, over the C host sources
 * (tests/native/Makefile); the tests fail on any sanitizer report (leaks included) or a wrong read-back.
 * usage: threads_host <learned model file>
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "libxstream_amd.h"

#define NTHREADS 8
#define NITER 100

static volatile int failures = 0, learned = 0;

static void* work(void* arg) {
  const int t = (int)(size_t)arg;
  int i;
  for (i = 0; i < NITER; ++i) {
    libsmm_acc_smm_config_t cfg, got;
    int src;
    char msg[64];
    const int m = 4 + t, n = 5 + i % 50, k = 6 + i % 7;
    memset(&cfg, 0, sizeof(cfg));
    cfg.variant = 0;
    cfg.tm = cfg.tn = 1;
    cfg.nt = 64;
    cfg.bs = 8 + t;
    cfg.flush = 2;
    cfg.s = 1000 + i;
    if (EXIT_SUCCESS == libsmm_acc_config_valid(dbcsr_type_real_8, m, n, k, &cfg)) {
      if (EXIT_SUCCESS != libsmm_acc_config_set(dbcsr_type_real_8, m, n, k, &cfg) ||
          EXIT_SUCCESS != libsmm_acc_config_get(dbcsr_type_real_8, m, n, k, &got, &src) ||
          libsmm_acc_config_tuned != src || got.bs != 8 + t || got.s != 1000 + i) {
        __atomic_add_fetch(&failures, 1, __ATOMIC_RELAXED);
      }
    }
    /* shared shapes: tuned rows and untuned ones that the learned slot (host inference) resolves */
    if (EXIT_SUCCESS != libsmm_acc_config_get(dbcsr_type_real_8, 23, 23, 23, &got, NULL) ||
        EXIT_SUCCESS != libsmm_acc_config_get(dbcsr_type_real_4, 17 + i % 10, 9 + t, 33, &got, &src) ||
        (libsmm_acc_config_learned == src && 0 == __atomic_add_fetch(&learned, 1, __ATOMIC_RELAXED)) ||
        EXIT_SUCCESS != libsmm_acc_config_heuristic(dbcsr_type_real_8, 30 + t, 31, 7 + i % 20, &got)) {
      __atomic_add_fetch(&failures, 1, __ATOMIC_RELAXED);
    }
    /* the calling thread's error state: set by an invalid call, read back by the same thread */
    if (EXIT_SUCCESS == libsmm_acc_config_set(dbcsr_type_real_8, 0, n, k, &cfg) ||
        0 == libxstream_error_consume(msg, sizeof(msg))) {
      __atomic_add_fetch(&failures, 1, __ATOMIC_RELAXED);
    }
  }
  return NULL;
}

int main(int argc, char* argv[]) {
  pthread_t threads[NTHREADS];
  libsmm_acc_mlp_t* model = NULL;
  int i;
  if (EXIT_SUCCESS != libsmm_acc_init()) return 2;
  if (1 < argc && (EXIT_SUCCESS != libsmm_acc_mlp_load(&model, argv[1]) ||
                   EXIT_SUCCESS != libsmm_acc_learned_set(model, 0.0))) {
    fprintf(stderr, "threads_host: cannot load the model %s\n", argv[1]);
    return 2;
  }
  for (i = 0; i < NTHREADS; ++i) pthread_create(threads + i, NULL, work, (void*)(size_t)i);
  for (i = 0; i < NTHREADS; ++i) pthread_join(threads[i], NULL);
  libsmm_acc_finalize();
  libsmm_acc_mlp_destroy(model); /* the caller keeps ownership of an installed model */
  printf("threads_host: %d threads x %d iterations, %d learned lookups, %d failures\n", NTHREADS, NITER, learned,
    failures);
  return 0 == failures ? 0 : 1;
}
