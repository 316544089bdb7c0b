/*
 * tune_multiply: exhaustive autotuner for the stack kernels (MI355X).
 *
 * Behaviour of the reference's tuner (samples/smm/tune_multiply.py:66-716, driven through
 * acc_bench.x and OPENCL_LIBSMM_SMM_* variables) without its search library: for every
 * (datatype, m, n, k) it measures every compiled kernel that fits the shape, times each
 * configuration with HIP events on the launch stream (like acc_bench.c:330-344: warm-up, then
 * repetitions over one synthetic stack), checks the fastest result against a host multiply
 * (CHECK=1 of tune_multiply.py:288, :639-646), and writes
 *   - one JSON line per measured configuration to the sweep file (resumable: shapes already
 *     marked done there are skipped, tune_multiply.py:656-671), and
 *   - the best configuration per shape to a CSV in the MI355X format
 *     DEVICE;TYPEID;M;N;K;S;GFLOPS;VARIANT;TM;TN;NT;BS;KT;FLUSH;SPLIT;BSS;SPLITS (load with
 *     OPENCL_LIBSMM_SMM_PARAMS).
 * After the sweep, bs and the long-run cut factor are refined around the winner: a refinement that is
 * faster at the tuned stack size replaces BS/SPLIT if it is not slower at 16 times that size, and is
 * kept as the short-stack pair BSS/SPLITS (used for stacks of at most S entries) otherwise.
 *
 * usage: tune_multiply [-t f64|f32] [-s stack] [-r reps] [-o best.csv] [-j sweep.jsonl]
 *                      [-d device] [-T seconds] [-l model.txt [-k topk]] [-x 1]
 *                      [-f triplets.txt | m n k [m n k ...]]
 * -x 1 only measures and records the sweep (training data for the learned predictor): no selection.
 * -T stops before the next shape once the time budget is spent (a resumed run continues there).
 * -l makes it learned-guided: only the model's topk (default 4) configurations are measured
 * (libsmm_acc_mlp_predict_topk) instead of every valid one.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "../../include/libxstream_amd.h"

typedef struct tune_opts_t {
  int dtype, stack_size, reps;
  const char *csv, *jsonl, *device;
  libsmm_acc_mlp_t* model; /* learned-guided: measure only the model's topk configurations */
  int topk;
  int sweep_only; /* -x: measure and record every configuration, choose nothing (training sweeps) */
} tune_opts_t;

static uint64_t splitmix64(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double uniform(uint64_t* state) { /* [-1, 1) */
  return 2.0 * (double)(splitmix64(state) >> 11) * (1.0 / 9007199254740992.0) - 1.0;
}

/* sorted-by-C stack: runs of 30 +- 4 entries over nc = S/30 C blocks, A/B pools of 2*nc blocks */
typedef struct tune_work_t {
  int m, n, k, stack_size, nc, na, nb, ts;
  int* stack;
  void *a, *b, *c;
} tune_work_t;

static int work_init(tune_work_t* w, int dtype, int m, int n, int k, int stack_size, uint64_t seed) {
  uint64_t st = seed;
  long i, count;
  int c = 0, left;
  memset(w, 0, sizeof(*w));
  w->m = m;
  w->n = n;
  w->k = k;
  w->stack_size = stack_size;
  w->ts = (dbcsr_type_real_8 == dtype ? 8 : 4);
  w->nc = (stack_size / 30 > 0 ? stack_size / 30 : 1);
  w->na = w->nb = 2 * w->nc;
  if ((double)w->na * m * k >= 2147483647.0 || (double)w->nc * m * n >= 2147483647.0) return EXIT_FAILURE;
  w->stack = (int*)malloc(sizeof(int) * 3 * (size_t)stack_size);
  w->a = malloc((size_t)w->ts * w->na * m * k);
  w->b = malloc((size_t)w->ts * w->nb * k * n);
  w->c = calloc((size_t)w->nc * m * n, (size_t)w->ts);
  if (NULL == w->stack || NULL == w->a || NULL == w->b || NULL == w->c) return EXIT_FAILURE;
  left = 26 + (int)(splitmix64(&st) % 9);
  for (i = 0; i < stack_size; ++i) {
    if (0 == left && c + 1 < w->nc) {
      ++c;
      left = 26 + (int)(splitmix64(&st) % 9);
    }
    if (0 < left) --left;
    w->stack[3 * i + 0] = (int)(splitmix64(&st) % (uint64_t)w->na) * m * k + 1;
    w->stack[3 * i + 1] = (int)(splitmix64(&st) % (uint64_t)w->nb) * k * n + 1;
    w->stack[3 * i + 2] = c * m * n + 1;
  }
  count = (long)w->na * m * k;
  for (i = 0; i < count; ++i) {
    if (8 == w->ts) ((double*)w->a)[i] = uniform(&st);
    else ((float*)w->a)[i] = (float)uniform(&st);
  }
  count = (long)w->nb * k * n; /* row-major k-by-n blocks: the layout the kernels read */
  for (i = 0; i < count; ++i) {
    if (8 == w->ts) ((double*)w->b)[i] = uniform(&st);
    else ((float*)w->b)[i] = (float)uniform(&st);
  }
  return EXIT_SUCCESS;
}

static void work_free(tune_work_t* w) {
  free(w->stack);
  free(w->a);
  free(w->b);
  free(w->c);
}

static double elem(const void* p, int ts, long i) { return 8 == ts ? ((const double*)p)[i] : ((const float*)p)[i]; }

/* host check of the first `nruns` C blocks: max |C - C_host| / max |C_host| (double accumulation) */
static double host_check(const tune_work_t* w, const void* cdev_copy, int nblocks) {
  const int m = w->m, n = w->n, k = w->k, mn = m * n;
  double* ref = (double*)calloc((size_t)nblocks * mn, sizeof(double));
  double maxref = 0, maxerr = 0;
  long s;
  int i;
  if (NULL == ref) return INFINITY;
  for (s = 0; s < w->stack_size; ++s) {
    const long a0 = w->stack[3 * s] - 1, b0 = w->stack[3 * s + 1] - 1, c0 = w->stack[3 * s + 2] - 1;
    int r, col, l;
    if (c0 / mn >= nblocks) continue;
    for (col = 0; col < n; ++col) {
      for (r = 0; r < m; ++r) {
        double sum = 0;
        for (l = 0; l < k; ++l) sum += elem(w->a, w->ts, a0 + (long)l * m + r) * elem(w->b, w->ts, b0 + (long)l * n + col);
        ref[c0 + (long)col * m + r] += sum;
      }
    }
  }
  for (i = 0; i < nblocks * mn; ++i) {
    const double d = fabs(elem(cdev_copy, w->ts, i) - ref[i]);
    if (fabs(ref[i]) > maxref) maxref = fabs(ref[i]);
    if (!(d <= maxerr)) maxerr = d;
  }
  free(ref);
  return maxerr / (maxref > 0 ? maxref : 1);
}

static int shape_done(const char* jsonl, int dtype, int m, int n, int k, int stack_size) {
  char line[1024], key[160];
  FILE* const f = (NULL != jsonl ? fopen(jsonl, "r") : NULL);
  int done = 0;
  if (NULL == f) return 0;
  snprintf(key, sizeof(key), "\"done\": 1, \"dtype\": %i, \"m\": %i, \"n\": %i, \"k\": %i, \"S\": %i,", dtype, m, n, k,
    stack_size);
  while (!done && NULL != fgets(line, sizeof(line), f)) done = (NULL != strstr(line, key));
  fclose(f);
  return done;
}

/* ms per launch of cfg: reps launches back to back between two events on the stream, after one untimed
   launch (acc_bench.c:330-344 times its repetitions the same way); -1 if the configuration does not run */
static double tune_time(const libsmm_acc_smm_config_t* cfg, const tune_work_t* w, const void* ds, const void* da,
  const void* db, void* dc, int m, int n, int k, int reps, void* stream) {
  const libsmm_acc_data_t dtype = (libsmm_acc_data_t)(8 == w->ts ? 3 : 1);
  void *e0 = NULL, *e1 = NULL;
  double ms = -1;
  int r, ok;
  if (EXIT_SUCCESS != libsmm_acc_process_config(cfg, (const int*)ds, w->stack_size, dtype, da, db, dc, m, n, k,
                        stream)) {
    (void)libxstream_error_consume(NULL, 0);
    return -1;
  }
  ok = (EXIT_SUCCESS == c_dbcsr_acc_event_create(&e0) && EXIT_SUCCESS == c_dbcsr_acc_event_create(&e1) &&
        EXIT_SUCCESS == c_dbcsr_acc_event_record(e0, stream));
  for (r = 0; r < reps && 0 != ok; ++r) {
    ok = (EXIT_SUCCESS == libsmm_acc_process_config(cfg, (const int*)ds, w->stack_size, dtype, da, db, dc, m, n, k,
                            stream));
  }
  if (0 != ok && EXIT_SUCCESS == c_dbcsr_acc_event_record(e1, stream) &&
      EXIT_SUCCESS == c_dbcsr_acc_event_synchronize(e1) &&
      EXIT_SUCCESS == libxstream_event_elapsed((libxstream_event_t*)e0, (libxstream_event_t*)e1, &ms) && 0 < ms) {
    ms /= (double)reps;
  }
  else {
    ms = -1;
    (void)libxstream_error_consume(NULL, 0);
  }
  if (NULL != e0) c_dbcsr_acc_event_destroy(e0);
  if (NULL != e1) c_dbcsr_acc_event_destroy(e1);
  return ms;
}

/* max normalised error of one launch of cfg from zeroed C, over the first C blocks */
static double tune_error(const libsmm_acc_smm_config_t* cfg, const tune_work_t* w, const void* ds, const void* da,
  const void* db, void* dc, int m, int n, int k, void* stream) {
  const int nblocks = (w->nc < 40 ? w->nc : 40);
  double err = 1;
  void* host = malloc((size_t)w->ts * nblocks * m * n);
  if (NULL == host) return err;
  c_dbcsr_acc_memset_zero(dc, 0, (size_t)w->ts * w->nc * m * n, stream);
  libsmm_acc_process_config(cfg, (const int*)ds, w->stack_size, (libsmm_acc_data_t)(8 == w->ts ? 3 : 1), da, db, dc, m, n, k,
    stream);
  c_dbcsr_acc_memcpy_d2h(dc, host, (size_t)w->ts * nblocks * m * n, stream);
  c_dbcsr_acc_stream_sync(stream);
  err = host_check(w, host, nblocks);
  free(host);
  return err;
}

#define TUNE_RERANK 6 /* fastest configurations of the sweep timed again, interleaved */
#define TUNE_ROUNDS 3 /* timing rounds of the re-ranking (median) */

static double median3(const double* t) {
  const double a = t[0], b = t[1], c = t[2];
  return (a < b) ? ((b < c) ? b : (a < c ? c : a)) : ((a < c) ? a : (b < c ? c : b));
}

static int tune_shape(const tune_opts_t* o, int m, int n, int k) {
  const int dtype = o->dtype;
  const char* const env_grouped = getenv("LIBXSTREAM_SMM_GROUPED");
  const int grouped = (NULL != env_grouped && 0 != atoi(env_grouped));
  tune_work_t w;
  void *da = NULL, *db = NULL, *dc = NULL, *ds = NULL, *stream = NULL;
  int cand[5 * 256], ncand, i, j, b, f, result = EXIT_SUCCESS, nmeasured = 0, ntodo = 0;
  static const int bss[] = {16, 32, 64, 128};
  static libsmm_acc_smm_config_t todo[2048];
  libsmm_acc_smm_config_t best, order[2048], top[TUNE_RERANK];
  double times[2048], flops, ts_s[TUNE_RERANK][TUNE_ROUNDS], ts_l[TUNE_RERANK][TUNE_ROUNDS];
  double med_s[TUNE_RERANK], med_l[TUNE_RERANK], score[TUNE_RERANK];
  tune_work_t wl;
  void *la = NULL, *lb = NULL, *lc = NULL, *ls = NULL;
  int have_long = 0, ntop = 0, r;
  FILE* sweep = NULL;
  if (shape_done(o->jsonl, dtype, m, n, k, o->stack_size)) {
    fprintf(stderr, "tune_multiply: %ix%ix%i already tuned (skipped)\n", m, n, k);
    return EXIT_SUCCESS;
  }
  if (EXIT_SUCCESS != work_init(&w, dtype, m, n, k, o->stack_size, 0x2507197525071975ull ^ (uint64_t)(m * 7919 + n * 104729 + k))) {
    work_free(&w);
    return EXIT_FAILURE;
  }
  flops = 2.0 * m * n * k * (double)w.stack_size;
  result |= c_dbcsr_acc_stream_create(&stream, "tune", -1);
  result |= c_dbcsr_acc_dev_mem_allocate(&da, (size_t)w.ts * w.na * m * k);
  result |= c_dbcsr_acc_dev_mem_allocate(&db, (size_t)w.ts * w.nb * k * n);
  result |= c_dbcsr_acc_dev_mem_allocate(&dc, (size_t)w.ts * w.nc * m * n);
  result |= c_dbcsr_acc_dev_mem_allocate(&ds, sizeof(int) * 3 * (size_t)w.stack_size);
  if (EXIT_SUCCESS == result) {
    result |= c_dbcsr_acc_memcpy_h2d(w.a, da, (size_t)w.ts * w.na * m * k, stream);
    result |= c_dbcsr_acc_memcpy_h2d(w.b, db, (size_t)w.ts * w.nb * k * n, stream);
    result |= c_dbcsr_acc_memcpy_h2d(w.stack, ds, sizeof(int) * 3 * (size_t)w.stack_size, stream);
    result |= c_dbcsr_acc_memset_zero(dc, 0, (size_t)w.ts * w.nc * m * n, stream);
    result |= c_dbcsr_acc_stream_sync(stream);
  }
  /* the configurations to measure: every valid one, or the learned model's topk */
  if (NULL != o->model) {
    if (EXIT_SUCCESS != libsmm_acc_mlp_predict_topk(o->model, (libsmm_acc_data_t)dtype, m, n, k,
                          (o->topk < 2048 ? o->topk : 2048), todo, NULL, &ntodo)) {
      ntodo = 0;
    }
  }
  else {
    ncand = libsmm_acc_kernel_candidates((libsmm_acc_data_t)dtype, m, n, k, cand, 256);
    for (i = 0; i < ncand && i < 256; ++i) {
      for (b = 0; b < 4; ++b) {
        for (f = 1; f <= 2; ++f) {
          libsmm_acc_smm_config_t cfg;
          memset(&cfg, 0, sizeof(cfg));
          cfg.variant = cand[5 * i];
          cfg.tm = cand[5 * i + 1];
          cfg.tn = cand[5 * i + 2];
          cfg.nt = cand[5 * i + 3];
          cfg.kt = cand[5 * i + 4];
          cfg.bs = bss[b];
          cfg.flush = f;
          if (EXIT_SUCCESS != libsmm_acc_config_valid((libsmm_acc_data_t)dtype, m, n, k, &cfg)) {
            if (0 != cfg.variant) continue;
            for (cfg.kt = (k > 8 ? (k - 1) / 8 * 8 : k - 1); 0 < cfg.kt; cfg.kt -= (cfg.kt > 8 ? 8 : 1)) {
              if (EXIT_SUCCESS == libsmm_acc_config_valid((libsmm_acc_data_t)dtype, m, n, k, &cfg)) break;
            }
            if (0 >= cfg.kt) continue;
          }
          if (ntodo < 2048) todo[ntodo++] = cfg;
        }
      }
    }
  }
  if (NULL != o->jsonl) sweep = fopen(o->jsonl, "a");
  for (i = 0; i < ntodo && EXIT_SUCCESS == result; ++i) {
    libsmm_acc_smm_config_t cfg = todo[i];
    const double ms = tune_time(&cfg, &w, ds, da, db, dc, m, n, k, o->reps, stream);
    if (0 >= ms) continue;
    cfg.gflops = flops / (ms * 1e6);
    if (nmeasured < 2048) {
      order[nmeasured] = cfg;
      times[nmeasured++] = ms;
    }
    if (NULL != sweep) {
      fprintf(sweep,
        "{\"dtype\": %i, \"m\": %i, \"n\": %i, \"k\": %i, \"S\": %i, \"variant\": %i, \"tm\": %i, \"tn\": %i, "
        "\"nt\": %i, \"bs\": %i, \"kt\": %i, \"flush\": %i, \"ms\": %.6f, \"gflops\": %.2f}\n",
        dtype, m, n, k, w.stack_size, cfg.variant, cfg.tm, cfg.tn, cfg.nt, cfg.bs, cfg.kt, cfg.flush, ms, cfg.gflops);
    }
  }
  if (0 != o->sweep_only) { /* the sweep file is the product: mark the shape done */
    if (NULL != sweep) {
      double gbest = 0;
      for (i = 0; i < nmeasured; ++i) gbest = (order[i].gflops > gbest ? order[i].gflops : gbest);
      fprintf(sweep, "{\"done\": 1, \"dtype\": %i, \"m\": %i, \"n\": %i, \"k\": %i, \"S\": %i, \"gflops\": %.2f}\n", dtype,
        m, n, k, w.stack_size, gbest);
      fclose(sweep);
    }
    printf("%ix%ix%i %s: %i configurations measured\n", m, n, k, 8 == w.ts ? "f64" : "f32", nmeasured);
    c_dbcsr_acc_dev_mem_deallocate(da);
    c_dbcsr_acc_dev_mem_deallocate(db);
    c_dbcsr_acc_dev_mem_deallocate(dc);
    c_dbcsr_acc_dev_mem_deallocate(ds);
    c_dbcsr_acc_stream_destroy(stream);
    work_free(&w);
    return result;
  }
  /* fastest first */
  for (i = 0; i < nmeasured; ++i) {
    for (j = i + 1; j < nmeasured; ++j) {
      if (times[j] < times[i]) {
        const double t = times[i];
        const libsmm_acc_smm_config_t c = order[i];
        times[i] = times[j];
        order[i] = order[j];
        times[j] = t;
        order[j] = c;
      }
    }
  }
  /* the long stack (16 x S entries, at most 500000): a row serves every longer stack as tuned */
  memset(&wl, 0, sizeof(wl));
  {
    const int s2 = (16 * w.stack_size < 500000 ? 16 * w.stack_size : 500000);
    have_long = (s2 > w.stack_size &&
                 EXIT_SUCCESS == work_init(&wl, dtype, m, n, k, s2, 0x5eed2507ull ^ (uint64_t)m) &&
                 EXIT_SUCCESS == c_dbcsr_acc_dev_mem_allocate(&la, (size_t)wl.ts * wl.na * m * k) &&
                 EXIT_SUCCESS == c_dbcsr_acc_dev_mem_allocate(&lb, (size_t)wl.ts * wl.nb * k * n) &&
                 EXIT_SUCCESS == c_dbcsr_acc_dev_mem_allocate(&lc, (size_t)wl.ts * wl.nc * m * n) &&
                 EXIT_SUCCESS == c_dbcsr_acc_dev_mem_allocate(&ls, sizeof(int) * 3 * (size_t)wl.stack_size) &&
                 EXIT_SUCCESS == c_dbcsr_acc_memcpy_h2d(wl.a, la, (size_t)wl.ts * wl.na * m * k, stream) &&
                 EXIT_SUCCESS == c_dbcsr_acc_memcpy_h2d(wl.b, lb, (size_t)wl.ts * wl.nb * k * n, stream) &&
                 EXIT_SUCCESS == c_dbcsr_acc_memcpy_h2d(wl.stack, ls, sizeof(int) * 3 * (size_t)wl.stack_size, stream) &&
                 EXIT_SUCCESS == c_dbcsr_acc_memset_zero(lc, 0, (size_t)wl.ts * wl.nc * m * n, stream) &&
                 EXIT_SUCCESS == c_dbcsr_acc_stream_sync(stream));
  }
  /* re-rank the TUNE_RERANK fastest eligible configurations: TUNE_ROUNDS rounds that time every candidate
     once at S and once on the long stack, interleaved (drift in clock or temperature hits them alike);
     score = geometric mean of the two median times, each relative to the best candidate at its size.
     Flush 1 (no atomics) launches as the atomic flush unless LIBXSTREAM_SMM_GROUPED is set: it is measured
     for the sweep file but cannot win. */
  for (i = 0; i < nmeasured && ntop < TUNE_RERANK; ++i) {
    if (1 == order[i].flush && 0 == grouped) continue;
    top[ntop++] = order[i];
  }
  for (r = 0; r < TUNE_ROUNDS && EXIT_SUCCESS == result; ++r) {
    for (i = 0; i < ntop; ++i) {
      ts_s[i][r] = tune_time(top + i, &w, ds, da, db, dc, m, n, k, o->reps, stream);
      ts_l[i][r] = (0 != have_long ? tune_time(top + i, &wl, ls, la, lb, lc, m, n, k, (o->reps + 3) / 4, stream) : 1);
    }
  }
  {
    double min_s = INFINITY, min_l = INFINITY;
    for (i = 0; i < ntop; ++i) {
      med_s[i] = median3(ts_s[i]);
      med_l[i] = median3(ts_l[i]);
      if (0 < med_s[i] && med_s[i] < min_s) min_s = med_s[i];
      if (0 < med_l[i] && med_l[i] < min_l) min_l = med_l[i];
    }
    for (i = 0; i < ntop; ++i) {
      score[i] = (0 < med_s[i] && 0 < med_l[i]) ? sqrt(med_s[i] / min_s * (med_l[i] / min_l)) : INFINITY;
      top[i].gflops = (0 < med_s[i] ? flops / (med_s[i] * 1e6) : 0);
    }
    for (i = 0; i < ntop; ++i) { /* by score */
      for (j = i + 1; j < ntop; ++j) {
        if (score[j] < score[i]) {
          const double t = score[i], u = med_s[i];
          const libsmm_acc_smm_config_t c = top[i];
          score[i] = score[j];
          med_s[i] = med_s[j];
          top[i] = top[j];
          score[j] = t;
          med_s[j] = u;
          top[j] = c;
        }
      }
    }
  }
  /* the first (by score) that passes the host check wins */
  memset(&best, 0, sizeof(best));
  for (i = 0; i < ntop && EXIT_SUCCESS == result; ++i) {
    const double err = tune_error(top + i, &w, ds, da, db, dc, m, n, k, stream);
    if (err <= (8 == w.ts ? 1e-13 : 1e-5)) {
      best = top[i];
      break;
    }
    fprintf(stderr, "tune_multiply: %ix%ix%i configuration %i failed the check (error %g)\n", m, n, k, i, err);
  }
  if (0 < best.nt && EXIT_SUCCESS == result) {
    /* refinement around the winner: entries per work-group and the long-run cut (C-runs longer than
       split x bs entries are cut into pieces); the sweep file keeps the plain configurations only */
    static const int rbs[] = {4, 8, 16, 32}, rsplit[] = {1, 2};
    libsmm_acc_smm_config_t cand_best = best;
    double t_best = tune_time(&best, &w, ds, da, db, dc, m, n, k, o->reps, stream);
    for (i = 0; i < 4 && 0 < t_best; ++i) {
      for (j = 0; j < 2; ++j) {
        libsmm_acc_smm_config_t cfg = best;
        double t;
        cfg.bs = rbs[i];
        cfg.split = rsplit[j];
        if (EXIT_SUCCESS != libsmm_acc_config_valid((libsmm_acc_data_t)dtype, m, n, k, &cfg)) continue;
        t = tune_time(&cfg, &w, ds, da, db, dc, m, n, k, o->reps, stream);
        if (0 < t && t < 0.97 * t_best) {
          cfg.gflops = flops / (t * 1e6);
          cand_best = cfg;
          t_best = t;
        }
      }
    }
    if (cand_best.bs != best.bs || cand_best.split != best.split) {
      /* a refinement must not lose on the long stack: bs is scaled down for shorter stacks only, so the
         configuration also serves stacks of many times the tuned size */
      int ok = 0, short_pair = 0;
      if (0 != have_long) {
        const double t_plain = tune_time(&best, &wl, ls, la, lb, lc, m, n, k, o->reps, stream);
        const double t_cand = tune_time(&cand_best, &wl, ls, la, lb, lc, m, n, k, o->reps, stream);
        ok = (0 < t_plain && 0 < t_cand && t_cand <= 1.01 * t_plain);
        if (0 == ok) { /* kept for stacks of at most S entries only */
          printf("%ix%ix%i: bs=%i split=%i is faster at S=%i but not at S=%i (%.4f vs %.4f ms): short stacks only\n",
            m, n, k, cand_best.bs, cand_best.split, w.stack_size, wl.stack_size, t_cand, t_plain);
          short_pair = 1;
        }
      }
      if (0 != ok || 0 != short_pair) {
        const double err = tune_error(&cand_best, &w, ds, da, db, dc, m, n, k, stream);
        if (err <= (8 == w.ts ? 1e-13 : 1e-5)) {
          if (0 != ok) best = cand_best;
          else {
            best.bs_short = cand_best.bs;
            best.split_short = cand_best.split;
            best.gflops = cand_best.gflops; /* the row's GFLOPS are measured at S */
          }
        }
      }
    }
  }
  if (0 != have_long || NULL != la || NULL != lb || NULL != lc || NULL != ls) {
    c_dbcsr_acc_dev_mem_deallocate(la);
    c_dbcsr_acc_dev_mem_deallocate(lb);
    c_dbcsr_acc_dev_mem_deallocate(lc);
    c_dbcsr_acc_dev_mem_deallocate(ls);
  }
  work_free(&wl);
  if (0 < best.nt) {
    best.s = w.stack_size; /* bs was chosen for this stack size */
    libsmm_acc_config_set((libsmm_acc_data_t)dtype, m, n, k, &best);
    printf("%ix%ix%i %s: %.1f GFLOPS/s (variant=%i tm=%i tn=%i nt=%i bs=%i kt=%i flush=%i split=%i bss=%i splits=%i, "
      "%i configurations)\n", m, n, k, 8 == w.ts ? "f64" : "f32", best.gflops, best.variant, best.tm, best.tn, best.nt,
      best.bs, best.kt, best.flush, best.split, best.bs_short, best.split_short, nmeasured);
    if (NULL != o->csv) {
      FILE* const csv = fopen(o->csv, "a+");
      if (NULL != csv) {
        fseek(csv, 0, SEEK_END);
        if (0 == ftell(csv)) fprintf(csv, "DEVICE;TYPEID;M;N;K;S;GFLOPS;VARIANT;TM;TN;NT;BS;KT;FLUSH;SPLIT;BSS;SPLITS\n");
        fprintf(csv, "%s;%i;%i;%i;%i;%i;%.1f;%i;%i;%i;%i;%i;%i;%i;%i;%i;%i\n", o->device, dtype, m, n, k,
          w.stack_size, best.gflops, best.variant, best.tm, best.tn, best.nt, best.bs, best.kt, best.flush, best.split,
          best.bs_short, best.split_short);
        fclose(csv);
      }
    }
    if (NULL != sweep) {
      fprintf(sweep, "{\"done\": 1, \"dtype\": %i, \"m\": %i, \"n\": %i, \"k\": %i, \"S\": %i, \"gflops\": %.2f}\n", dtype,
        m, n, k, w.stack_size, best.gflops);
    }
  }
  else {
    fprintf(stderr, "tune_multiply: %ix%ix%i: no configuration passed\n", m, n, k);
    result = EXIT_FAILURE;
  }
  if (NULL != sweep) fclose(sweep);
  c_dbcsr_acc_dev_mem_deallocate(da);
  c_dbcsr_acc_dev_mem_deallocate(db);
  c_dbcsr_acc_dev_mem_deallocate(dc);
  c_dbcsr_acc_dev_mem_deallocate(ds);
  c_dbcsr_acc_stream_destroy(stream);
  work_free(&w);
  return result;
}

static double seconds_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int main(int argc, char* argv[]) {
  tune_opts_t o;
  int i, result, ndevices = 0, ntuned = 0;
  FILE* triplets = NULL;
  const char* model_path = NULL;
  double budget = 0;
  const double start = seconds_now();
  o.dtype = dbcsr_type_real_8;
  o.stack_size = 30000;
  o.reps = 20;
  o.csv = NULL;
  o.jsonl = NULL;
  o.device = "MI355X";
  o.model = NULL;
  o.topk = 0;
  o.sweep_only = 0;
  for (i = 1; i < argc && '-' == argv[i][0] && '\0' != argv[i][1] && i + 1 < argc; i += 2) {
    const char* const v = argv[i + 1];
    switch (argv[i][1]) {
      case 't': o.dtype = (0 == strcmp(v, "f32") ? dbcsr_type_real_4 : dbcsr_type_real_8); break;
      case 's': o.stack_size = atoi(v); break;
      case 'r': o.reps = atoi(v); break;
      case 'o': o.csv = v; break;
      case 'j': o.jsonl = v; break;
      case 'd': o.device = v; break;
      case 'f': triplets = fopen(v, "r"); break;
      case 'T': budget = atof(v); break;
      case 'l': model_path = v; break;
      case 'k': o.topk = atoi(v); break;
      case 'x': o.sweep_only = atoi(v); break;
      default: fprintf(stderr, "tune_multiply: unknown option %s\n", argv[i]); return EXIT_FAILURE;
    }
  }
  if (0 >= o.stack_size || 0 >= o.reps) return EXIT_FAILURE;
  result = libsmm_acc_init();
  if (EXIT_SUCCESS == result) result = c_dbcsr_acc_get_ndevices(&ndevices);
  if (EXIT_SUCCESS != result || 0 >= ndevices) {
    fprintf(stderr, "tune_multiply: no device\n");
    return EXIT_FAILURE;
  }
  if (NULL != model_path) {
    if (EXIT_SUCCESS != libsmm_acc_mlp_load(&o.model, model_path)) {
      fprintf(stderr, "tune_multiply: cannot load the model %s\n", model_path);
      return EXIT_FAILURE;
    }
    if (0 >= o.topk) o.topk = 4;
  }
  if (NULL != triplets) {
    int m, n, k;
    while (EXIT_SUCCESS == result && 3 == fscanf(triplets, "%i %i %i", &m, &n, &k)) {
      if (0 < budget && seconds_now() - start > budget) {
        printf("tune_multiply: time budget of %.0f s reached after %i shapes\n", budget, ntuned);
        break;
      }
      result = tune_shape(&o, m, n, k);
      ++ntuned;
    }
    fclose(triplets);
  }
  for (; i + 2 < argc + 0 && EXIT_SUCCESS == result; i += 3) {
    result = tune_shape(&o, atoi(argv[i]), atoi(argv[i + 1]), atoi(argv[i + 2]));
  }
  if (NULL != o.model) libsmm_acc_mlp_destroy(o.model);
  libsmm_acc_finalize();
  return result;
}
