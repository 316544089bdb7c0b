/*
 * ACC LIBSMM host layer: libsmm_acc_init/finalize/is_thread_safe/process/transpose.
 *
 * Follows the reference's host contract (samples/smm/smm_acc.c:43-604, smm_kernel.c:82-528,
 * smm_trans.c:48-228): suitability gate with the -1 host-fallback result, per-shape configuration
 * registry, tuned parameters from a CSV file (OPENCL_LIBSMM_SMM_PARAMS), environment overrides,
 * asynchronous launch on the caller's stream with flops/bytes stated per launch. The kernels are
 * compiled ahead of time for gfx950, so there is no JIT step; a configuration selects one of the
 * compiled kernel variants.
 */
#if !defined(_GNU_SOURCE)
#  define _GNU_SOURCE /* dladdr */
#endif
#include "lxs_internal.h"
#include <ctype.h>
#include <dlfcn.h>
#include <string.h>
#include <strings.h>

#define LXS_SMM_MAXDIM_DEFAULT 80
/* stack size of the autotuner's sweeps: the built-in and learned configurations are chosen for it */
#define LXS_SMM_S 30000
/* smallest bs the scaling for shorter stacks starts from (smm_kernel.c:49-51, 466-467) */
#define LXS_SMM_BS_MIN 32

typedef struct lxs_smm_slot_t {
  int type, m, n, k, used, source;
  int devmatch; /* 1: a CSV row of the device that best matches the active one */
  libsmm_acc_smm_config_t config;
} lxs_smm_slot_t;

static pthread_mutex_t lxs_smm_lock = PTHREAD_MUTEX_INITIALIZER;
static lxs_smm_slot_t* lxs_smm_slots = NULL;
static int lxs_smm_cap = 0, lxs_smm_count = 0;
static int lxs_smm_ninit = 0;

static int lxs_typesize(int datatype) {
  return (dbcsr_type_real_8 == datatype ? 8 : (dbcsr_type_real_4 == datatype ? 4 : 0));
}

/* registry ------------------------------------------------------------------------------------ */

static unsigned int lxs_smm_hash(int type, int m, int n, int k) {
  unsigned int h = 2166136261u;
  const int v[4] = {type, m, n, k};
  int i;
  for (i = 0; i < 4; ++i) {
    h ^= (unsigned int)v[i];
    h *= 16777619u;
  }
  return h;
}

static lxs_smm_slot_t* lxs_smm_lookup(int type, int m, int n, int k, int insert) { /* lock held */
  unsigned int i, mask;
  if (0 == lxs_smm_cap || (0 != insert && 2 * (lxs_smm_count + 1) > lxs_smm_cap)) {
    const int cap = (0 < lxs_smm_cap ? 2 * lxs_smm_cap : 256);
    lxs_smm_slot_t* const slots = (lxs_smm_slot_t*)calloc((size_t)cap, sizeof(lxs_smm_slot_t));
    lxs_smm_slot_t* const old = lxs_smm_slots;
    const int oldcap = lxs_smm_cap;
    int j;
    if (NULL == slots) return NULL;
    lxs_smm_slots = slots;
    lxs_smm_cap = cap;
    lxs_smm_count = 0;
    for (j = 0; j < oldcap; ++j) {
      if (0 != old[j].used) {
        lxs_smm_slot_t* const s = lxs_smm_lookup(old[j].type, old[j].m, old[j].n, old[j].k, 1);
        *s = old[j];
      }
    }
    free(old);
  }
  mask = (unsigned int)lxs_smm_cap - 1;
  for (i = lxs_smm_hash(type, m, n, k) & mask;; i = (i + 1) & mask) {
    lxs_smm_slot_t* const s = lxs_smm_slots + i;
    if (0 == s->used) {
      if (0 == insert) return NULL;
      memset(s, 0, sizeof(*s));
      s->used = 1;
      s->type = type;
      s->m = m;
      s->n = n;
      s->k = k;
      ++lxs_smm_count;
      return s;
    }
    if (s->type == type && s->m == m && s->n == n && s->k == k) return s;
  }
}

/* configurations ------------------------------------------------------------------------------ */

/* LDS/thread feasibility of a configuration for a shape */
LIBXSTREAM_API int libsmm_acc_config_valid(libsmm_acc_data_t datatype, int m, int n, int k,
  const libsmm_acc_smm_config_t* config) {
  lxs_smm_args_t args;
  if (NULL == config || 0 >= m || 0 >= n || 0 >= k || 0 == lxs_typesize(datatype)) return EXIT_FAILURE;
  if (0 > config->variant || 4 < config->variant || 0 >= config->tm || 0 >= config->tn || 0 >= config->nt ||
      0 > config->bs || 0 > config->kt || 0 > config->split || 0 > config->bs_short || 0 > config->split_short) {
    return EXIT_FAILURE;
  }
  if (0 == lxs_smm_available((int)datatype, config->variant, m, n, k, config->tm, config->tn, config->nt,
             0 != config->variant ? config->kt : 0)) {
    return EXIT_FAILURE;
  }
  if (1 == config->variant) return EXIT_SUCCESS; /* registers only */
  memset(&args, 0, sizeof(args));
  args.dtype = (int)datatype;
  args.variant = config->variant;
  args.m = m;
  args.n = n;
  args.k = k;
  args.tm = config->tm;
  args.tn = config->tn;
  args.nt = config->nt;
  args.kt = config->kt;
  return (0 != lxs_smm_launchable(&args)) ? EXIT_SUCCESS : EXIT_FAILURE;
}

/* largest k-slice that keeps the staged blocks of all groups within the LDS limit */
static int lxs_smm_fit_kt(int datatype, int m, int n, int k, libsmm_acc_smm_config_t* config) {
  int kt = (0 < config->kt && config->kt < k) ? config->kt : k;
  for (; 0 < kt; kt = (kt > 8 ? kt - 8 : kt - 1)) {
    config->kt = kt;
    if (EXIT_SUCCESS == libsmm_acc_config_valid((libsmm_acc_data_t)datatype, m, n, k, config)) {
      if (kt == k) config->kt = 0;
      return EXIT_SUCCESS;
    }
  }
  return EXIT_FAILURE;
}

/**
 * Built-in choice when nothing is tuned. FP64 blocks up to 16x16 go to the entry-parallel kernel on 4x4 blocks
 * that split k (variant 4; 4 wavefronts per run, 32 stack entries per work-group, a kernel compiled for the
 * shape first): on ten such shapes it ran 0.81-1.0 of the best variant-4 configuration, itself 1.0-2.2x
 * faster than the best on 16x16 tiles. Other C blocks of up to 16 tiles of 16x16 (m, n <= 64) go to the
 * entry-parallel matrix-core kernel (variant 2) for exactly that tile grid, with the most wavefronts per
 * run compiled for it (8 up to 4 tiles, 4 up to 8, 2 up to 16; a kernel compiled for the shape first) and
 * 32 (FP64) or 16 (FP32) stack entries per work-group: over the 419 such shapes of the autotuner's
 * 500-triplet sweeps this reaches 0.94 (FP64) and 0.92 (FP32) of each shape's best on average.
 * Larger blocks: the matrix-core kernel with one wavefront per run when at least 30% of the covered
 * area is useful (m*n >= 0.3 * 256 * ti * tj), a kernel compiled for exactly this shape first, else the
 * runtime-shape kernel for exactly that tile grid. Otherwise the vector kernel: 256 threads, the square
 * microtile (up to 5x5) that keeps the most threads busy with at most 16 groups of threads sharing a
 * C-run's entries (runs are tens of entries long), ties to the smaller tile; 64 threads for blocks too
 * small for that. Stack entries per work-group there: 32 (matrix cores) or 64 (vector kernel).
 */
static int lxs_smm_default(int datatype, int m, int n, int k, libsmm_acc_smm_config_t* config) {
  const int ti = (m + 15) / 16, tj = (n + 15) / 16;
  int q[5], i, t, count, cand[5 * 256], best_t = 0, best_nt = 256;
  double best_util = 0;
  memset(config, 0, sizeof(*config));
  if (dbcsr_type_real_8 == datatype && 16 >= m && 16 >= n) {
    const int bi = (m + 3) / 4, bj = (n + 3) / 4;
    count = lxs_smm_candidates(datatype, m, n, k, cand, 256);
    for (i = 0; i < count && i < 256; ++i) {
      const int* const c = cand + 5 * i;
      if (4 == c[0] && bi == c[1] && bj == c[2] && 256 == c[3]) {
        config->variant = 4;
        config->tm = bi;
        config->tn = bj;
        config->nt = 256;
        config->kt = c[4];
        config->bs = 32;
        config->s = LXS_SMM_S;
        if (EXIT_SUCCESS == libsmm_acc_config_valid((libsmm_acc_data_t)datatype, m, n, k, config)) return EXIT_SUCCESS;
        memset(config, 0, sizeof(*config));
        break;
      }
    }
  }
  if (16 >= ti * tj) {
    const int nt = 64 * (4 >= ti * tj ? 8 : (8 >= ti * tj ? 4 : 2));
    count = lxs_smm_candidates(datatype, m, n, k, cand, 256); /* kernels compiled for the shape come first */
    for (i = 0; i < count && i < 256; ++i) {
      const int* const c = cand + 5 * i;
      if (2 == c[0] && ti == c[1] && tj == c[2] && nt == c[3]) {
        config->variant = 2;
        config->tm = ti;
        config->tn = tj;
        config->nt = nt;
        config->kt = c[4];
        config->bs = (dbcsr_type_real_8 == datatype ? 32 : 16);
        config->s = LXS_SMM_S;
        if (EXIT_SUCCESS == libsmm_acc_config_valid((libsmm_acc_data_t)datatype, m, n, k, config)) return EXIT_SUCCESS;
        memset(config, 0, sizeof(*config));
        break;
      }
    }
  }
  if ((double)m * n >= 0.3 * 256.0 * ti * tj) {
    int found = (0 <= lxs_smm_specialised(datatype, 1, m, n, k, q));
    count = lxs_smm_variants(datatype, cand, 256);
    for (i = 0; i < count && i < 256 && !found; ++i) {
      const int* const c = cand + 5 * i;
      if (1 == c[0] && ti == c[1] && tj == c[2]) {
        memcpy(q, c, sizeof(q));
        found = 1;
      }
    }
    if (found) {
      config->variant = 1;
      config->tm = q[1];
      config->tn = q[2];
      config->nt = q[3];
      config->kt = q[4];
      config->bs = 32;
      config->s = LXS_SMM_S;
      return EXIT_SUCCESS;
    }
  }
  for (i = 0; i < 2 && 0 == best_t; ++i) { /* 256 threads, else 64 threads for tiny blocks */
    const int nt = (0 == i ? 256 : 64);
    for (t = 1; t <= 5; ++t) {
      const int ntiles = ((m + t - 1) / t) * ((n + t - 1) / t);
      if (ntiles <= nt && nt / ntiles <= 16) {
        const double util = (double)((nt / ntiles) * ntiles) / nt;
        if (util > best_util + 1e-9) {
          best_util = util;
          best_t = t;
          best_nt = nt;
        }
      }
    }
  }
  config->variant = 0;
  config->bs = 64;
  config->s = LXS_SMM_S;
  if (0 == best_t && m * n <= 64) { /* tiny blocks: one element per thread, 64 threads */
    best_t = 1;
    best_nt = 64;
  }
  if (0 < best_t) {
    config->tm = config->tn = best_t;
    config->nt = best_nt;
    if (EXIT_SUCCESS == lxs_smm_fit_kt(datatype, m, n, k, config)) return EXIT_SUCCESS;
  }
  config->tm = config->tn = 4; /* up to 88 x 88 with 512 threads */
  config->nt = 512;
  config->kt = 0;
  if (EXIT_SUCCESS == lxs_smm_fit_kt(datatype, m, n, k, config)) return EXIT_SUCCESS;
  memset(config, 0, sizeof(*config));
  return EXIT_FAILURE;
}

/* 2-character keys are suppressed when OPENCL_LIBSMM_SMM_PARAMS=0 (smm_acc.c:494-524) */
static const char* lxs_smm_getenv(const char* key) {
  char name[64];
  const char* const params = getenv("OPENCL_LIBSMM_SMM_PARAMS");
  if (2 == strlen(key) && NULL != params && '0' == *params) return NULL;
  snprintf(name, sizeof(name), "OPENCL_LIBSMM_SMM_%s", key);
  {
    const char* const value = getenv(name);
    return (NULL != value && '\0' != *value) ? value : NULL;
  }
}

static int lxs_smm_env(libsmm_acc_smm_config_t* config) {
  static const char* const keys[] = {"BS", "TM", "TN", "NT", "KT", "FL", "VA"};
  int i, any = 0;
  for (i = 0; i < 7; ++i) {
    const char* const value = lxs_smm_getenv(keys[i]);
    if (NULL != value) {
      const int v = atoi(value);
      switch (i) {
        case 0: config->bs = v; break;
        case 1: config->tm = v; break;
        case 2: config->tn = v; break;
        case 3: config->nt = v; break;
        case 4: config->kt = v; break;
        case 5: config->flush = v; break;
        default: config->variant = v;
      }
      any = 1;
    }
  }
  return any;
}

static int lxs_smm_register(int datatype, int m, int n, int k, const libsmm_acc_smm_config_t* config, int source,
  int keep_best, int devmatch);

static void lxs_smm_resolve(int datatype, int m, int n, int k, libsmm_acc_smm_config_t* config, int* source) {
  const lxs_smm_slot_t* slot;
  int src = libsmm_acc_config_default;
  pthread_mutex_lock(&lxs_smm_lock);
  slot = lxs_smm_lookup(datatype, m, n, k, 0);
  if (NULL != slot) {
    *config = slot->config;
    src = slot->source;
  }
  pthread_mutex_unlock(&lxs_smm_lock);
  if (NULL == slot) {
    if (EXIT_SUCCESS == lxs_learned_predict(datatype, m, n, k, config)) { /* cached for the next lookup */
      src = libsmm_acc_config_learned;
      config->s = LXS_SMM_S; /* the model learned from sweeps at this stack size */
      (void)lxs_smm_register(datatype, m, n, k, config, libsmm_acc_config_learned, 0, 0);
    }
    else {
      lxs_smm_default(datatype, m, n, k, config);
    }
  }
  {
    libsmm_acc_smm_config_t env = *config;
    if (0 != lxs_smm_env(&env)) {
      if (EXIT_SUCCESS == libsmm_acc_config_valid((libsmm_acc_data_t)datatype, m, n, k, &env)) {
        *config = env;
        src = libsmm_acc_config_env;
      }
      else if (0 != lxs_cfg.verbosity) {
        fprintf(stderr, "WARN LIBSMM: OPENCL_LIBSMM_SMM_* settings are not valid for %ix%ix%i (ignored)\n", m, n, k);
      }
    }
  }
  if (NULL != source) *source = src;
}

LIBXSTREAM_API int libsmm_acc_config_heuristic(libsmm_acc_data_t datatype, int m, int n, int k,
  libsmm_acc_smm_config_t* config) {
  if (NULL == config || 0 == lxs_typesize(datatype) || 0 >= m || 0 >= n || 0 >= k) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_config_heuristic: invalid argument");
  }
  return lxs_smm_default((int)datatype, m, n, k, config);
}

LIBXSTREAM_API int libsmm_acc_config_get(libsmm_acc_data_t datatype, int m, int n, int k, libsmm_acc_smm_config_t* config,
  int* source) {
  if (NULL == config || 0 == lxs_typesize(datatype) || 0 >= m || 0 >= n || 0 >= k) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_config_get: invalid argument");
  }
  lxs_smm_resolve((int)datatype, m, n, k, config, source);
  return (0 < config->nt ? EXIT_SUCCESS : EXIT_FAILURE);
}

/* keep_best = 0 replaces any entry. Otherwise a row of the device matching the active one (devmatch) wins
   over rows of other devices whatever their speed, and among rows of equal standing a duplicate replaces
   the entry only when it is faster (smm_acc.c:117-119, :204-235) */
static int lxs_smm_register(int datatype, int m, int n, int k, const libsmm_acc_smm_config_t* config, int source,
  int keep_best, int devmatch) {
  lxs_smm_slot_t* slot;
  int result = EXIT_SUCCESS;
  pthread_mutex_lock(&lxs_smm_lock);
  slot = lxs_smm_lookup(datatype, m, n, k, 1);
  if (NULL != slot) {
    if (0 == keep_best || 0 == slot->config.nt || slot->devmatch < devmatch ||
        (slot->devmatch == devmatch && slot->config.gflops < config->gflops)) {
      slot->config = *config;
      slot->source = source;
      slot->devmatch = devmatch;
    }
  }
  else {
    result = EXIT_FAILURE;
  }
  pthread_mutex_unlock(&lxs_smm_lock);
  return result;
}

LIBXSTREAM_API int libsmm_acc_config_set(libsmm_acc_data_t datatype, int m, int n, int k,
  const libsmm_acc_smm_config_t* config) {
  if (EXIT_SUCCESS != libsmm_acc_config_valid(datatype, m, n, k, config)) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_config_set: configuration not valid for this shape");
  }
  return lxs_smm_register((int)datatype, m, n, k, config, libsmm_acc_config_tuned, 0, 0);
}

LIBXSTREAM_API int libsmm_acc_kernel_variants(libsmm_acc_data_t datatype, int* quintuples, int cap) {
  return lxs_smm_variants((int)datatype, quintuples, cap);
}

LIBXSTREAM_API int libsmm_acc_kernel_candidates(libsmm_acc_data_t datatype, int m, int n, int k, int* quintuples,
  int cap) {
  return lxs_smm_candidates((int)datatype, m, n, k, quintuples, cap);
}

/* tuned-parameter CSV -------------------------------------------------------------------------- */

/* CSV column separators; spaces are part of a value (device names contain them) */
static const char lxs_delims[] = ";,\t\r\n";

LIBXSTREAM_API int libsmm_acc_read_ref_params(const char* line, int with_device, int with_perf, char* device,
  size_t device_size, int key[4], int params[16], double s_gflops[2]) {
  /* positional columns as in samples/smm/smm_params.c:99-243 */
  char buffer[1024], *save = NULL, *tok;
  int column = 0, nkey = 0, nparams = 0, nperf = 0;
  if (NULL == line || NULL == key || NULL == params) return EXIT_FAILURE;
  snprintf(buffer, sizeof(buffer), "%s", line);
  memset(params, 0, sizeof(int) * 16);
  if (NULL != s_gflops) s_gflops[0] = s_gflops[1] = 0;
  for (tok = strtok_r(buffer, lxs_delims, &save); NULL != tok; tok = strtok_r(NULL, lxs_delims, &save)) {
    char* end = NULL;
    while (' ' == *tok) ++tok;
    if (0 == column && 0 != with_device) {
      if (NULL != device && 0 < device_size) snprintf(device, device_size, "%s", tok);
      ++column;
      continue;
    }
    if ('\0' == *tok) return EXIT_FAILURE; /* empty column */
    if (4 > nkey) {
      const long v = strtol(tok, &end, 10);
      while (' ' == *end) ++end;
      if ('\0' != *end || (0 < nkey && 0 >= v)) return EXIT_FAILURE;
      key[nkey++] = (int)v;
    }
    else if (0 != with_perf && 2 > nperf) {
      const double v = strtod(tok, &end);
      while (' ' == *end) ++end;
      if ('\0' != *end || 0 > v) return EXIT_FAILURE;
      if (NULL != s_gflops) s_gflops[nperf] = v;
      ++nperf;
    }
    else if (16 > nparams) {
      const long v = strtol(tok, &end, 10);
      while (' ' == *end) ++end;
      if ('\0' != *end) return EXIT_FAILURE;
      params[nparams++] = (int)v;
    }
    else {
      return EXIT_FAILURE; /* too many columns */
    }
    ++column;
  }
  if (4 != nkey || (0 != with_perf && 2 != nperf) || 15 > nparams) return EXIT_FAILURE;
  return (dbcsr_type_real_8 == key[0] || dbcsr_type_real_4 == key[0]) ? EXIT_SUCCESS : EXIT_FAILURE;
}

static int lxs_column(char names[][16], int ncolumns, const char* name) {
  int i;
  for (i = 0; i < ncolumns; ++i) {
    if (0 == strcasecmp(names[i], name)) return i;
  }
  return -1;
}

/* lower-case alphanumeric words of a device name, without the "[0x...]" device id of the reference's rows */
static int lxs_device_words(const char* name, char words[][32], int cap) {
  int n = 0;
  while (NULL != name && '\0' != *name && n < cap) {
    int len = 0;
    while ('\0' != *name && !isalnum((unsigned char)*name)) ++name;
    while (isalnum((unsigned char)*name)) {
      if (len < 31) words[n][len++] = (char)tolower((unsigned char)*name);
      ++name;
    }
    words[n][len] = '\0';
    if (0 < len && !('0' == words[n][0] && 'x' == words[n][1])) ++n;
  }
  return n;
}

/* share of a row's device words found in the active device's name (0 .. 1) */
static double lxs_device_score(const char* row, char active[][32], int nactive) {
  char words[16][32];
  const int n = lxs_device_words(row, words, 16);
  int i, j, hits = 0;
  for (i = 0; i < n; ++i) {
    for (j = 0; j < nactive && 0 != strcmp(words[i], active[j]); ++j);
    hits += (j < nactive);
  }
  return 0 < n ? (double)hits / n : 0.0;
}

/* value of the given column of a CSV row ("" if the row is shorter) */
static void lxs_row_device(const char* line, int column, char* device, size_t size) {
  int i;
  for (i = 0; i < column && '\0' != *line; ++i) {
    line += strcspn(line, lxs_delims);
    if ('\0' != *line) ++line;
  }
  snprintf(device, size, "%.*s", (int)strcspn(line, lxs_delims), line);
}

LIBXSTREAM_API int libsmm_acc_params_load(const char* filename) {
  return libsmm_acc_params_load_device(filename, NULL);
}

LIBXSTREAM_API int libsmm_acc_params_load_device(const char* filename, const char* device) {
  char line[1024], names[32][16], best[256] = "", active[16][32];
  int ncolumns = 0, nrows = 0, native, with_device, with_perf, nactive, devcol;
  long body;
  FILE* const file = (NULL != filename ? fopen(filename, "r") : NULL);
  if (NULL == file) return -1;
  if (NULL == fgets(line, sizeof(line), file)) {
    fclose(file);
    return -1;
  }
  { /* header decides the format: MI355X columns (TM, NT, ...) or the reference's (BM, BN, ...) */
    char *save = NULL, *tok;
    for (tok = strtok_r(line, lxs_delims, &save); NULL != tok && ncolumns < 32; tok = strtok_r(NULL, lxs_delims, &save)) {
      snprintf(names[ncolumns++], sizeof(names[0]), "%s", tok);
    }
  }
  native = (0 <= lxs_column(names, ncolumns, "TM") && 0 <= lxs_column(names, ncolumns, "NT"));
  with_device = (0 <= lxs_column(names, ncolumns, "DEVICE"));
  with_perf = (0 <= lxs_column(names, ncolumns, "GFLOPS"));
  devcol = lxs_column(names, ncolumns, "DEVICE");
  if (0 > lxs_column(names, ncolumns, "TYPEID") || (0 != with_device && 0 == native && 0 != devcol)) {
    fclose(file);
    return -1;
  }
  /* rows of several devices: those of the device whose name best matches the active one take precedence
     (smm_acc.c:150-235); the other devices' rows fill the remaining shapes, the fastest row per shape */
  nactive = lxs_device_words(NULL != device ? device : lxs_cfg.devname, active, 16);
  body = ftell(file);
  if (0 != with_device && 0 < nactive && 0 <= body) {
    double best_score = 0;
    while (NULL != fgets(line, sizeof(line), file)) {
      char row[256];
      double score;
      lxs_row_device(line, devcol, row, sizeof(row));
      score = lxs_device_score(row, active, nactive);
      if (best_score < score) {
        best_score = score;
        snprintf(best, sizeof(best), "%s", row);
      }
    }
    if (0 != fseek(file, body, SEEK_SET)) best[0] = '\0';
    if (0 != lxs_cfg.verbosity && '\0' != best[0]) {
      fprintf(stderr, "INFO LIBSMM: tuned parameters of \"%s\" take precedence\n", best);
    }
  }
  while (NULL != fgets(line, sizeof(line), file)) {
    libsmm_acc_smm_config_t config;
    int key[4], ok = 0, devmatch = 0;
    if ('\0' == line[strspn(line, lxs_delims)]) continue; /* blank line */
    if ('\0' != best[0]) {
      char row[256];
      lxs_row_device(line, devcol, row, sizeof(row));
      devmatch = (0 == strcmp(row, best));
    }
    memset(&config, 0, sizeof(config));
    if (0 != native) {
      char *save = NULL, *tok, values[32][64];
      int i = 0;
      for (tok = strtok_r(line, lxs_delims, &save); NULL != tok && i < 32; tok = strtok_r(NULL, lxs_delims, &save)) {
        snprintf(values[i++], sizeof(values[0]), "%s", tok);
      }
      if (i == ncolumns) {
#define LXS_GETI(NAME) (0 <= lxs_column(names, ncolumns, NAME) ? atoi(values[lxs_column(names, ncolumns, NAME)]) : 0)
        key[0] = LXS_GETI("TYPEID");
        key[1] = LXS_GETI("M");
        key[2] = LXS_GETI("N");
        key[3] = LXS_GETI("K");
        config.variant = LXS_GETI("VARIANT");
        config.tm = LXS_GETI("TM");
        config.tn = LXS_GETI("TN");
        config.nt = LXS_GETI("NT");
        config.bs = LXS_GETI("BS");
        config.kt = LXS_GETI("KT");
        config.flush = LXS_GETI("FLUSH");
        config.s = LXS_GETI("S");
        config.split = LXS_GETI("SPLIT");
        config.bs_short = LXS_GETI("BSS");
        config.split_short = LXS_GETI("SPLITS");
#undef LXS_GETI
        if (0 != with_perf) config.gflops = atof(values[lxs_column(names, ncolumns, "GFLOPS")]);
        ok = (EXIT_SUCCESS == libsmm_acc_config_valid((libsmm_acc_data_t)key[0], key[1], key[2], key[3], &config));
      }
    }
    else {
      int params[16];
      double perf[2];
      if (EXIT_SUCCESS == libsmm_acc_read_ref_params(line, with_device, with_perf, NULL, 0, key, params, perf)) {
        /* only BS (stack entries per work-group) carries over to these kernels */
        ok = (EXIT_SUCCESS == lxs_smm_default(key[0], key[1], key[2], key[3], &config));
        config.bs = (0 < params[0] ? params[0] : config.bs);
        config.s = (int)perf[0];
        config.gflops = perf[1];
      }
    }
    if (0 != ok) {
      if (EXIT_SUCCESS != lxs_smm_register(key[0], key[1], key[2], key[3], &config, libsmm_acc_config_tuned, 1,
                             devmatch)) {
        break;
      }
      ++nrows;
    }
    else if (0 != lxs_cfg.verbosity) {
      fprintf(stderr, "WARN LIBSMM: skipped tuned-parameter row: %s", line);
    }
  }
  fclose(file);
  return nrows;
}

LIBXSTREAM_API int libsmm_acc_params_save(const char* filename, const char* device) {
  FILE* const file = (NULL != filename ? fopen(filename, "w") : NULL);
  int i, nrows = 0;
  if (NULL == file) return -1;
  fprintf(file, "DEVICE;TYPEID;M;N;K;S;GFLOPS;VARIANT;TM;TN;NT;BS;KT;FLUSH;SPLIT;BSS;SPLITS\n");
  pthread_mutex_lock(&lxs_smm_lock);
  for (i = 0; i < lxs_smm_cap; ++i) {
    const lxs_smm_slot_t* const s = lxs_smm_slots + i;
    if (0 == s->used) continue;
    fprintf(file, "%s;%i;%i;%i;%i;%i;%.1f;%i;%i;%i;%i;%i;%i;%i;%i;%i;%i\n", NULL != device ? device : "MI355X", s->type,
      s->m, s->n, s->k, s->config.s, s->config.gflops, s->config.variant, s->config.tm, s->config.tn, s->config.nt,
      s->config.bs, s->config.kt, s->config.flush, s->config.split, s->config.bs_short, s->config.split_short);
    ++nrows;
  }
  pthread_mutex_unlock(&lxs_smm_lock);
  fclose(file);
  return nrows;
}

/* directory of this library (tuned parameters ship next to it) */
/* path of a file shipped in params/ next to the loaded library */
static int lxs_smm_default_params(char* path, size_t size, const char* file) {
  Dl_info info;
  if (0 != dladdr((void*)&lxs_smm_default_params, &info) && NULL != info.dli_fname) {
    const char* const slash = strrchr(info.dli_fname, '/');
    const int len = (NULL != slash ? (int)(slash - info.dli_fname) : 0);
    snprintf(path, size, "%.*s%sparams/%s", len, info.dli_fname, 0 < len ? "/" : "", file);
    return EXIT_SUCCESS;
  }
  return EXIT_FAILURE;
}

int lxs_smm_init(void) {
  const char* const env = getenv("OPENCL_LIBSMM_SMM_PARAMS");
  char path[4096];
  if (NULL == env) {
    if (EXIT_SUCCESS == lxs_smm_default_params(path, sizeof(path), "tune_multiply_MI355X.csv")) {
      (void)libsmm_acc_params_load(path);
    }
  }
  else if ('0' != *env && '\0' != *env) {
    if (0 > libsmm_acc_params_load(env)) {
      /* not a file: "TYPEID M N K VARIANT TM TN NT BS KT FLUSH" as one configuration */
      libsmm_acc_smm_config_t config;
      int key[4];
      memset(&config, 0, sizeof(config));
      if (11 == sscanf(env, "%i %i %i %i %i %i %i %i %i %i %i", key, key + 1, key + 2, key + 3, &config.variant,
                   &config.tm, &config.tn, &config.nt, &config.bs, &config.kt, &config.flush)) {
        if (EXIT_SUCCESS != libsmm_acc_config_set((libsmm_acc_data_t)key[0], key[1], key[2], key[3], &config)) {
          return EXIT_FAILURE;
        }
      }
      else if (0 != lxs_cfg.verbosity) {
        fprintf(stderr, "WARN LIBSMM: failed to open parameter file!\n");
      }
    }
  }
  { /* the shipped predictor serves shapes without a tuned configuration (LIBXSTREAM_LEARNED=0: off) */
    const char* const learned = getenv("LIBXSTREAM_LEARNED");
    if ((NULL == learned || '0' != *learned) &&
        EXIT_SUCCESS == lxs_smm_default_params(path, sizeof(path), "mlp_MI355X.txt")) {
      FILE* const f = fopen(path, "r");
      if (NULL != f) {
        fclose(f);
        lxs_learned_load_default(path);
      }
    }
  }
  return EXIT_SUCCESS;
}

void lxs_smm_finalize(void) {
  lxs_learned_finalize();
  pthread_mutex_lock(&lxs_smm_lock);
  free(lxs_smm_slots);
  lxs_smm_slots = NULL;
  lxs_smm_cap = lxs_smm_count = 0;
  pthread_mutex_unlock(&lxs_smm_lock);
}

/* public entry points -------------------------------------------------------------------------- */

LIBXSTREAM_API int libsmm_acc_init(void) {
  int result = EXIT_SUCCESS;
  pthread_mutex_lock(&lxs_smm_lock);
  ++lxs_smm_ninit;
  pthread_mutex_unlock(&lxs_smm_lock);
  if (1 == lxs_smm_ninit) {
    result = libxstream_init();
    if (EXIT_SUCCESS == result) result = lxs_smm_init();
  }
  return result;
}

LIBXSTREAM_API int libsmm_acc_finalize(void) {
  int result = EXIT_SUCCESS, last;
  pthread_mutex_lock(&lxs_smm_lock);
  last = (0 < lxs_smm_ninit && 0 == --lxs_smm_ninit);
  pthread_mutex_unlock(&lxs_smm_lock);
  if (0 != last) {
    lxs_smm_finalize();
    result = libxstream_finalize();
  }
  return result;
}

LIBXSTREAM_API c_dbcsr_acc_bool_t libsmm_acc_is_thread_safe(void) { return 1; }

LIBXSTREAM_API c_dbcsr_acc_bool_t libsmm_acc_process_suitable(c_dbcsr_acc_bool_t def_mnk, libsmm_acc_data_t datatype,
  int stack_size, int m_max, int n_max, int k_max, int max_kernel_dim) {
  const long mn = (long)m_max * (long)n_max;
  c_dbcsr_acc_bool_t result = 0;
  if (0 < mn && 0 < k_max && 0 < stack_size && 0 != def_mnk && mn <= (long)max_kernel_dim * max_kernel_dim) {
    result = (dbcsr_type_real_8 == datatype || dbcsr_type_real_4 == datatype);
  }
  if (0 == result && 2 <= lxs_cfg.verbosity) {
    fprintf(stderr, "INFO ACC/LIBSMM: SMM-kernel {%i,%i,%i,%i} ss=%i %s\n", (int)datatype, m_max, n_max, k_max,
      stack_size, mn <= (long)max_kernel_dim * max_kernel_dim ? (0 != def_mnk ? "is ignored" : "is inhomogeneous")
                                                              : "is too large");
  }
  return result;
}

/* A C-run is summed by the work-group where it starts, unless it is longer than tsplit = split x bs
   entries (the configuration's split, else LIBXSTREAM_SMM_SPLIT, default 2; LIBXSTREAM_SMM_SPLIT=0 never
   cuts): then it is cut at work-group boundaries and its pieces are added with atomics, so one long run
   (many k-blocks into one C block) does not leave a single work-group working alone. Floor: in a stack of
   at least LXS_SMM_SPLIT_FILL entries (enough C-runs to occupy the device), runs of up to 64 entries are
   never cut: each is summed in one work-group in a fixed order and written once, so its C block gets the
   same bits on every call and for every bs. Shorter stacks may cut any run longer than tsplit, which
   spreads their few runs over more work-groups. LIBXSTREAM_SMM_SPLIT_MIN=n sets the floor to n entries
   for every stack (1: no floor). */
int lxs_smm_tsplit(int bs, int split, int stack_size) {
  const int factor = (0 < split ? split : lxs_cfg.split);
  const int t = factor * (1 < bs ? bs : 1);
  const int floor = (0 < lxs_cfg.split_min ? lxs_cfg.split_min : (LXS_SMM_SPLIT_FILL <= stack_size ? 64 : 1));
  return (0 < lxs_cfg.split ? (floor < t ? t : floor) : 0);
}

/* the launch arguments of one stack under a configuration (bs scaling, short-stack pair, cut length) */
static void lxs_smm_fill(lxs_smm_args_t* out, const libsmm_acc_smm_config_t* config, const int* dev_param_stack,
  int stack_size, int datatype, const void* a, const void* b, void* c, int m, int n, int k, int param_format,
  int grouped) {
  lxs_smm_args_t args;
  int split;
  memset(&args, 0, sizeof(args));
  args.c = c;
  args.a = a;
  args.b = b;
  args.stack = dev_param_stack;
  args.stack_size = stack_size;
  args.m = m;
  args.n = n;
  args.k = k;
  args.dtype = datatype;
  args.variant = config->variant;
  args.tm = config->tm;
  args.tn = config->tn;
  args.nt = config->nt;
  args.bs = config->bs;
  args.kt = config->kt;
  split = config->split;
  if (0 < config->bs_short && 0 < config->s && stack_size <= config->s) { /* the short-stack pair */
    args.bs = config->bs_short;
    split = config->split_short;
  }
  /* a stack shorter than the one bs was chosen for spreads over more work-groups (smm_kernel.c:463-473) */
  if (1 < args.bs && 1 < config->s && stack_size < config->s) {
    const long bs0 = args.bs, cbs = (bs0 > LXS_SMM_BS_MIN ? bs0 : LXS_SMM_BS_MIN);
    const long bs = ((long)stack_size * cbs + config->s - 2) / (config->s - 1);
    args.bs = (int)(bs < bs0 ? (1 > bs ? 1 : bs) : bs0);
  }
  /* flush: atomic add, which is correct for any stack order (multiply.cl:657,687) and exact-once per
     element for a C-run that is not cut; flush 1 (plain read-modify-write) only when the caller has
     guaranteed (grouped) that each C block of the stack is one contiguous run, because two work-groups
     owning separate runs of one C block would otherwise lose updates */
  args.atomic = (1 == config->flush && 0 != grouped) ? 0 : 1;
  args.pfmt = param_format;
  args.tsplit = lxs_smm_tsplit(args.bs, split, stack_size);
  *out = args;
}

static int lxs_smm_run(const libsmm_acc_smm_config_t* config, const int* dev_param_stack, int stack_size,
  int datatype, const void* a, const void* b, void* c, int m, int n, int k, void* stream, int param_format,
  int grouped) {
  const hipStream_t hs = lxs_stream_hip((const libxstream_stream_t*)stream);
  lxs_smm_args_t args;
  void* token;
  int result = lxs_ensure_device();
  if (EXIT_SUCCESS != result) return result;
  lxs_smm_fill(&args, config, dev_param_stack, stack_size, datatype, a, b, c, m, n, k, param_format, grouped);
  token = lxs_profile_begin(hs);
  result = lxs_smm_launch(&args, (void*)hs);
  lxs_profile_end(token, hs, "smm", 2.0 * stack_size * m * n * k,
    (double)stack_size * lxs_typesize(datatype) * ((double)m * k + (double)k * n + 2.0 * m * n));
  return result;
}

/* one stack with its resolved configuration on a HIP stream, without profile records (the batch's
   per-stack mode, also under graph capture); the result is a hipError_t or an LXS_E* code */
int lxs_smm_enqueue(int datatype, const int* dev_param_stack, int stack_size, const void* a, const void* b, void* c,
  int m, int n, int k, void* hip_stream) {
  libsmm_acc_smm_config_t config;
  lxs_smm_args_t args;
  if (0 >= stack_size) return 0;
  lxs_smm_resolve(datatype, m, n, k, &config, NULL);
  lxs_smm_fill(&args, &config, dev_param_stack, stack_size, datatype, a, b, c, m, n, k, 0, lxs_cfg.grouped);
  return lxs_smm_launch(&args, hip_stream);
}

LIBXSTREAM_API int libsmm_acc_process_config(const libsmm_acc_smm_config_t* config, const int* dev_param_stack,
  int stack_size, libsmm_acc_data_t datatype, const void* dev_a_data, const void* dev_b_data, void* dev_c_data, int m,
  int n, int k, void* stream) {
  int result;
  if (0 == stack_size) return EXIT_SUCCESS;
  if (0 > stack_size || NULL == dev_param_stack || NULL == dev_a_data || NULL == dev_b_data || NULL == dev_c_data) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_process_config: invalid argument");
  }
  if (EXIT_SUCCESS != libsmm_acc_config_valid(datatype, m, n, k, config)) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_process_config: configuration not valid for this shape");
  }
  /* an explicit configuration with flush 1 is the caller's statement that the stack is grouped by C */
  result = lxs_smm_run(config, dev_param_stack, stack_size, (int)datatype, dev_a_data, dev_b_data, dev_c_data, m, n, k,
    stream, 0, 1);
  return (0 <= result ? lxs_hip((hipError_t)result, "stack kernel launch") : lxs_error(result, "stack kernel launch"));
}

/* the DBM launch function registered by opencl_libsmm_acc_set_dbm_launch_fn (NULL: none) */
static opencl_libsmm_acc_dbm_launch_fn_t lxs_dbm_launch_fn;

LIBXSTREAM_API void opencl_libsmm_acc_set_dbm_launch_fn(opencl_libsmm_acc_dbm_launch_fn_t launch_fn) {
  __atomic_store_n(&lxs_dbm_launch_fn, launch_fn, __ATOMIC_RELEASE);
}

/* libsmm_acc_process over records of either format (param_format 0: DBCSR triplets, else DBM tasks) */
static int lxs_smm_process(const int* dev_param_stack, int stack_size, libsmm_acc_data_t datatype,
  const void* dev_a_data, const void* dev_b_data, void* dev_c_data, int m_max, int n_max, int k_max, int max_kernel_dim,
  c_dbcsr_acc_bool_t def_mnk, void* stack_stream, int param_format) {
  libsmm_acc_smm_config_t config;
  int result;
  if (0 == libsmm_acc_process_suitable(def_mnk, datatype, stack_size, m_max, n_max, k_max, max_kernel_dim)) {
    return (0 < stack_size ? -1 : EXIT_SUCCESS); /* smm_kernel.c:524-526 */
  }
  if (NULL == dev_param_stack || NULL == dev_a_data || NULL == dev_b_data || NULL == dev_c_data) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_process: NULL buffer");
  }
  lxs_smm_resolve((int)datatype, m_max, n_max, k_max, &config, NULL);
  if (0 >= config.nt) return -1; /* no compiled kernel covers this shape: host fallback */
  result = lxs_smm_run(&config, dev_param_stack, stack_size, (int)datatype, dev_a_data, dev_b_data, dev_c_data, m_max,
    n_max, k_max, stack_stream, param_format, lxs_cfg.grouped);
  if (LXS_ENOVARIANT == result || LXS_ELIMIT == result) {
    if (EXIT_SUCCESS != lxs_smm_default((int)datatype, m_max, n_max, k_max, &config)) return -1;
    result = lxs_smm_run(&config, dev_param_stack, stack_size, (int)datatype, dev_a_data, dev_b_data, dev_c_data,
      m_max, n_max, k_max, stack_stream, param_format, lxs_cfg.grouped);
    if (0 > result) return -1;
  }
  return lxs_hip((hipError_t)result, "stack kernel launch");
}

LIBXSTREAM_API int libsmm_acc_process(const int* host_param_stack, const int* dev_param_stack, int stack_size,
  libsmm_acc_data_t datatype, const void* dev_a_data, const void* dev_b_data, void* dev_c_data, int m_max, int n_max,
  int k_max, int max_kernel_dim, c_dbcsr_acc_bool_t def_mnk, void* stack_stream, void* c_stream) {
  const opencl_libsmm_acc_dbm_launch_fn_t dbm = __atomic_load_n(&lxs_dbm_launch_fn, __ATOMIC_ACQUIRE);
  (void)c_stream;
  /* smm_acc.c:579-585: FP64 stacks of one shape go to a registered DBM launch function, the shape packed
     into its param_format argument (8 bits per extent) */
  if (dbcsr_type_real_8 == datatype && 0 != def_mnk && NULL != dbm && 0 <= m_max && 0 <= n_max && 0 <= k_max &&
      m_max < (1 << (LIBSMM_ACC_PFORMAT - 1)) && n_max < (1 << (LIBSMM_ACC_PFORMAT - 1)) &&
      k_max < (1 << (LIBSMM_ACC_PFORMAT - 1)))
  {
    return dbm(stack_stream, 1.0, stack_size,
      m_max | (n_max << LIBSMM_ACC_PFORMAT) | (k_max << (2 * LIBSMM_ACC_PFORMAT)), host_param_stack, dev_param_stack,
      (const double*)dev_a_data, (const double*)dev_b_data, (double*)dev_c_data);
  }
  return lxs_smm_process(dev_param_stack, stack_size, datatype, dev_a_data, dev_b_data, dev_c_data, m_max, n_max, k_max,
    max_kernel_dim, def_mnk, stack_stream, 0);
}

LIBXSTREAM_API int opencl_libsmm_acc_process(const int* host_param_stack, const int* dev_param_stack, int stack_size,
  libsmm_acc_data_t datatype, const void* dev_a_data, const void* dev_b_data, void* dev_c_data, int m_max, int n_max,
  int k_max, int max_kernel_dim, c_dbcsr_acc_bool_t def_mnk, void* stream, void* c_stream, int param_format,
  void* event) {
  int result;
  (void)host_param_stack;
  (void)c_stream;
  result = lxs_smm_process(dev_param_stack, stack_size, datatype, dev_a_data, dev_b_data, dev_c_data, m_max, n_max,
    k_max, max_kernel_dim, def_mnk, stream, param_format);
  if (EXIT_SUCCESS == result && NULL != event) { /* completes when the launch has */
    result = libxstream_event_record((libxstream_event_t*)event, (libxstream_stream_t*)stream);
  }
  return result;
}

LIBXSTREAM_API int libsmm_acc_transpose(const int* dev_trs_stack, int offset, int stack_size, void* dev_data,
  libsmm_acc_data_t datatype, int m, int n, int max_kernel_dim, void* stream) {
  const long mn = (long)m * (long)n;
  const hipStream_t hs = lxs_stream_hip((const libxstream_stream_t*)stream);
  void* token;
  int result;
  if (0 >= m || 0 >= n || 0 > stack_size || 0 > offset) return lxs_error(EXIT_FAILURE, "libsmm_acc_transpose: invalid argument");
  /* smm_trans.c:54-68: nothing to do for these cases, which is not an error */
  if (0 == stack_size || 1 == mn || 0 == lxs_typesize(datatype) || mn > (long)max_kernel_dim * max_kernel_dim) {
    return EXIT_SUCCESS;
  }
  if (NULL == dev_trs_stack || NULL == dev_data) return lxs_error(EXIT_FAILURE, "libsmm_acc_transpose: NULL buffer");
  result = lxs_ensure_device();
  if (EXIT_SUCCESS != result) return result;
  token = lxs_profile_begin(hs);
  result = lxs_trans_launch(dev_trs_stack, offset, stack_size, dev_data, (int)datatype, m, n, (void*)hs);
  lxs_profile_end(token, hs, "trans", 0, 2.0 * stack_size * (double)mn * lxs_typesize(datatype));
  return (0 <= result ? lxs_hip((hipError_t)result, "transpose launch") : lxs_error(result, "transpose launch"));
}

LIBXSTREAM_API int c_calculate_norms(const double* mat, int nblks, const int* offsets, const int* nelems, float* norms,
  void* stream_ptr) {
  (void)mat;
  (void)nblks;
  (void)offsets;
  (void)nelems;
  (void)norms;
  (void)stream_ptr;
  return -1;
}
