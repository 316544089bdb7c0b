/*
 * Batched stack multiply: stacks of different (m, n, k) and lengths in one persistent kernel
 * launch, optionally replayed from a captured hipGraph. Each stack has the semantics of
 * libsmm_acc_process (samples/smm/acc_libsmm.h:45-47); the batch has no reference counterpart.
 *
 * Plan (built once on the host): every stack is cut into chunks of run-aligned entries, sized so
 * that a chunk moves about as many operand elements as 32 entries of 23x23x23; blocks that need
 * more than 4 output tiles of 16x16 get one work-group task per chunk (its 4 wavefronts split the
 * tiles), smaller blocks are packed 4 chunks per task (one per wavefront). Expensive tasks come
 * first; the persistent grid strides over the task list.
 */
#include "lxs_internal.h"
#include <string.h>

#define LXS_BATCH_CHUNK_ELEMENTS (32 * 2 * 23 * 23)
#define LXS_BATCH_LANES 4 /* side streams of the concurrent per-stack mode */

struct libsmm_acc_batch_t {
  int dtype, nstacks, ntasks, nblocks;
  lxs_batch_stack_t* desc;  /* device */
  lxs_batch_task_t* tasks;  /* device */
  hipStream_t capture;
  hipGraph_t graph;
  hipGraphExec_t exec;
  double flops, bytes;
  int per_stack;                 /* launch mode, libsmm_acc_batch_params_t */
  libsmm_acc_batch_stack_t* host; /* the stacks as given (per-stack mode) */
  /* per_stack 2: the launches spread over LXS_BATCH_LANES side streams that fork from the caller's stream
     and join it again (one graph with independent branches under replay); lane[i]: stack i's stream */
  hipStream_t side[LXS_BATCH_LANES];
  hipEvent_t fork, join[LXS_BATCH_LANES];
  int* lane;
};

typedef struct lxs_chunk_t {
  int stack, chunk;
  double cost;
} lxs_chunk_t;

static int lxs_chunk_cmp(const void* x, const void* y) {
  const double a = ((const lxs_chunk_t*)x)->cost, b = ((const lxs_chunk_t*)y)->cost;
  return (a < b) - (a > b); /* descending */
}

LIBXSTREAM_API int libsmm_acc_batch_create(libsmm_acc_data_t datatype, const libsmm_acc_batch_stack_t* stacks,
  int nstacks, libsmm_acc_batch_t** batch) {
  return libsmm_acc_batch_create_ex(datatype, stacks, nstacks, NULL, batch);
}

/* everything a batch holds besides its graph */
static void lxs_batch_release(libsmm_acc_batch_t* b) {
  int i;
  if (NULL != b->capture) (void)hipStreamDestroy(b->capture);
  for (i = 0; i < LXS_BATCH_LANES; ++i) {
    if (NULL != b->side[i]) (void)hipStreamDestroy(b->side[i]);
    if (NULL != b->join[i]) (void)hipEventDestroy(b->join[i]);
  }
  if (NULL != b->fork) (void)hipEventDestroy(b->fork);
  if (NULL != b->desc) (void)hipFree(b->desc);
  if (NULL != b->tasks) (void)hipFree(b->tasks);
  free(b->lane);
  free(b->host);
  free(b);
}

LIBXSTREAM_API int libsmm_acc_batch_create_ex(libsmm_acc_data_t datatype, const libsmm_acc_batch_stack_t* stacks,
  int nstacks, const libsmm_acc_batch_params_t* params, libsmm_acc_batch_t** batch) {
  const int chunk_elements = (NULL != params && 0 < params->chunk_elements ? params->chunk_elements
                                                                            : LXS_BATCH_CHUNK_ELEMENTS);
  const int blocks_per_cu = (NULL != params && 0 < params->blocks_per_cu ? params->blocks_per_cu : 8);
  const int ts = (dbcsr_type_real_8 == datatype ? 8 : (dbcsr_type_real_4 == datatype ? 4 : 0));
  lxs_batch_stack_t* desc = NULL;
  lxs_batch_task_t* tasks = NULL;
  lxs_chunk_t *big = NULL, *small = NULL;
  libsmm_acc_batch_t* b = NULL;
  int i, nbig = 0, nsmall = 0, ntasks = 0, result;
  if (NULL == batch || NULL == stacks || 0 >= nstacks || 0 == ts) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_batch_create: invalid argument");
  }
  *batch = NULL;
  result = lxs_ensure_device();
  if (EXIT_SUCCESS != result) return result;
  desc = (lxs_batch_stack_t*)calloc((size_t)nstacks, sizeof(*desc));
  b = (libsmm_acc_batch_t*)calloc(1, sizeof(*b));
  if (NULL == desc || NULL == b) {
    free(desc);
    free(b);
    return lxs_error(EXIT_FAILURE, "libsmm_acc_batch_create: out of memory");
  }
  for (i = 0; i < nstacks; ++i) { /* validate, size the chunks, count them */
    const libsmm_acc_batch_stack_t* const s = stacks + i;
    lxs_batch_stack_t* const d = desc + i;
    if (0 >= s->m || 64 < s->m || 0 >= s->n || 64 < s->n || 0 >= s->k || 0 > s->stack_size ||
        (0 < s->stack_size && (NULL == s->dev_param_stack || NULL == s->dev_a_data || NULL == s->dev_b_data ||
                                 NULL == s->dev_c_data))) {
      free(desc);
      free(b);
      return lxs_error(EXIT_FAILURE, "libsmm_acc_batch_create: stack with m, n outside 1..64 or NULL buffers");
    }
    d->a = s->dev_a_data;
    d->b = s->dev_b_data;
    d->c = s->dev_c_data;
    d->stack = s->dev_param_stack;
    d->stack_size = s->stack_size;
    d->m = s->m;
    d->n = s->n;
    d->k = s->k;
    d->ti = (s->m + 15) / 16;
    d->tj = (s->n + 15) / 16;
    d->bs = chunk_elements / (s->m * s->k + s->k * s->n);
    if (d->bs < 8) d->bs = 8;
    if (d->bs > 512) d->bs = 512;
    d->tsplit = lxs_smm_tsplit(d->bs, 0, s->stack_size);
    if (4 < d->ti * d->tj) nbig += (s->stack_size + d->bs - 1) / d->bs;
    else nsmall += (s->stack_size + d->bs - 1) / d->bs;
    b->flops += 2.0 * s->m * s->n * s->k * (double)s->stack_size;
    b->bytes += (double)ts * s->stack_size * ((double)s->m * s->k + (double)s->k * s->n + 2.0 * s->m * s->n);
  }
  big = (lxs_chunk_t*)malloc(sizeof(lxs_chunk_t) * (size_t)(nbig + 1));
  small = (lxs_chunk_t*)malloc(sizeof(lxs_chunk_t) * (size_t)(nsmall + 1));
  tasks = (lxs_batch_task_t*)malloc(sizeof(lxs_batch_task_t) * (size_t)(nbig + (nsmall + 3) / 4 + 1));
  if (NULL == big || NULL == small || NULL == tasks) {
    result = lxs_error(EXIT_FAILURE, "libsmm_acc_batch_create: out of memory");
  }
  if (EXIT_SUCCESS == result) {
    int c, j;
    nbig = nsmall = 0;
    for (i = 0; i < nstacks; ++i) {
      const lxs_batch_stack_t* const d = desc + i;
      const int nchunks = (d->stack_size + d->bs - 1) / d->bs;
      /* cost of a chunk ~ the operand bytes it moves; a team chunk has 4 wavefronts on it */
      const double cost = (double)d->bs * (d->m * d->k + d->k * d->n) / (4 < d->ti * d->tj ? 4.0 : 1.0);
      for (c = 0; c < nchunks; ++c) {
        lxs_chunk_t* const x = (4 < d->ti * d->tj ? big + nbig++ : small + nsmall++);
        x->stack = i;
        x->chunk = c;
        x->cost = cost;
      }
    }
    qsort(big, (size_t)nbig, sizeof(*big), lxs_chunk_cmp);
    qsort(small, (size_t)nsmall, sizeof(*small), lxs_chunk_cmp);
    for (i = 0; i < nbig; ++i) { /* team tasks first: they are the expensive ones */
      lxs_batch_task_t* const t = tasks + ntasks++;
      memset(t, 0, sizeof(*t));
      t->team = 1;
      t->stack[0] = big[i].stack;
      t->chunk[0] = big[i].chunk;
      for (j = 1; j < 4; ++j) t->stack[j] = -1;
    }
    for (i = 0; i < nsmall; i += 4) {
      lxs_batch_task_t* const t = tasks + ntasks++;
      memset(t, 0, sizeof(*t));
      for (j = 0; j < 4; ++j) {
        t->stack[j] = (i + j < nsmall ? small[i + j].stack : -1);
        t->chunk[j] = (i + j < nsmall ? small[i + j].chunk : 0);
      }
    }
  }
  if (EXIT_SUCCESS == result) {
    b->dtype = (int)datatype;
    b->nstacks = nstacks;
    b->ntasks = ntasks;
    b->per_stack = (NULL != params && 0 < params->per_stack ? (2 <= params->per_stack ? 2 : 1) : 0);
    b->host = (libsmm_acc_batch_stack_t*)malloc(sizeof(*stacks) * (size_t)nstacks);
    if (NULL == b->host) result = lxs_error(EXIT_FAILURE, "libsmm_acc_batch_create: out of memory");
    else memcpy(b->host, stacks, sizeof(*stacks) * (size_t)nstacks);
    for (i = 0; i < nstacks && 0 != b->per_stack && EXIT_SUCCESS == result; ++i) {
      /* resolve (and cache) every shape's configuration now, so that no lookup or prediction runs
         while libsmm_acc_batch_replay captures the launches */
      libsmm_acc_smm_config_t cfg;
      result = libsmm_acc_config_get(datatype, stacks[i].m, stacks[i].n, stacks[i].k, &cfg, NULL);
    }
  }
  if (EXIT_SUCCESS == result && 2 == b->per_stack) { /* lanes: stacks by flops, each to the least loaded */
    double load[LXS_BATCH_LANES] = {0};
    int* const order = (int*)malloc(sizeof(int) * (size_t)nstacks);
    b->lane = (int*)malloc(sizeof(int) * (size_t)nstacks);
    if (NULL == order || NULL == b->lane) result = lxs_error(EXIT_FAILURE, "libsmm_acc_batch_create: out of memory");
    for (i = 0; i < nstacks && EXIT_SUCCESS == result; ++i) order[i] = i;
    for (i = 0; i < nstacks && EXIT_SUCCESS == result; ++i) { /* descending flops (insertion sort) */
      const int x = order[i];
      const double fx = (double)stacks[x].m * stacks[x].n * stacks[x].k * stacks[x].stack_size;
      int j = i;
      while (0 < j && (double)stacks[order[j - 1]].m * stacks[order[j - 1]].n * stacks[order[j - 1]].k *
                          stacks[order[j - 1]].stack_size < fx) {
        order[j] = order[j - 1];
        --j;
      }
      order[j] = x;
    }
    for (i = 0; i < nstacks && EXIT_SUCCESS == result; ++i) {
      const int x = order[i];
      int l, best = 0;
      for (l = 1; l < LXS_BATCH_LANES; ++l) best = (load[l] < load[best] ? l : best);
      b->lane[x] = best;
      load[best] += (double)stacks[x].m * stacks[x].n * stacks[x].k * stacks[x].stack_size;
    }
    free(order);
    for (i = 0; i < LXS_BATCH_LANES && EXIT_SUCCESS == result; ++i) {
      result = lxs_hip(hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking), "batch: side stream");
      if (EXIT_SUCCESS == result) {
        result = lxs_hip(hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming), "batch: join event");
      }
    }
    if (EXIT_SUCCESS == result) result = lxs_hip(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "batch: fork event");
  }
  if (EXIT_SUCCESS == result) {
    b->nblocks = (0 < lxs_cfg.ncus ? blocks_per_cu * lxs_cfg.ncus : 256 * blocks_per_cu);
    if (b->nblocks > ntasks) b->nblocks = ntasks;
    result = lxs_hip(hipMalloc((void**)&b->desc, sizeof(*desc) * (size_t)nstacks), "hipMalloc batch");
  }
  if (EXIT_SUCCESS == result && 0 < ntasks) {
    result = lxs_hip(hipMalloc((void**)&b->tasks, sizeof(*tasks) * (size_t)ntasks), "hipMalloc batch");
  }
  if (EXIT_SUCCESS == result) {
    result = lxs_hip(hipMemcpy(b->desc, desc, sizeof(*desc) * (size_t)nstacks, hipMemcpyHostToDevice), "batch upload");
  }
  if (EXIT_SUCCESS == result && 0 < ntasks) {
    result = lxs_hip(hipMemcpy(b->tasks, tasks, sizeof(*tasks) * (size_t)ntasks, hipMemcpyHostToDevice), "batch upload");
  }
  free(big);
  free(small);
  free(tasks);
  free(desc);
  if (EXIT_SUCCESS == result) {
    *batch = b;
  }
  else {
    lxs_batch_release(b);
  }
  return result;
}

static int lxs_batch_enqueue(libsmm_acc_batch_t* b, hipStream_t hs) {
  int i, code = 0;
  if (0 == b->per_stack) {
    code = lxs_batch_launch(b->dtype, b->desc, b->tasks, b->ntasks, b->nblocks, (void*)hs);
  }
  else if (1 == b->per_stack) {
    for (i = 0; i < b->nstacks && 0 == code; ++i) {
      const libsmm_acc_batch_stack_t* const s = b->host + i;
      code = lxs_smm_enqueue(b->dtype, s->dev_param_stack, s->stack_size, s->dev_a_data, s->dev_b_data, s->dev_c_data,
        s->m, s->n, s->k, (void*)hs);
    }
  }
  else { /* fork to the side streams, one launch per stack on its lane, join (the C flush is atomic, so
            stacks that share a C pool may overlap) */
    int l;
    code = (int)hipEventRecord(b->fork, hs);
    for (l = 0; l < LXS_BATCH_LANES && 0 == code; ++l) code = (int)hipStreamWaitEvent(b->side[l], b->fork, 0);
    for (i = 0; i < b->nstacks && 0 == code; ++i) {
      const libsmm_acc_batch_stack_t* const s = b->host + i;
      code = lxs_smm_enqueue(b->dtype, s->dev_param_stack, s->stack_size, s->dev_a_data, s->dev_b_data, s->dev_c_data,
        s->m, s->n, s->k, (void*)b->side[b->lane[i]]);
    }
    for (l = 0; l < LXS_BATCH_LANES && 0 == code; ++l) {
      code = (int)hipEventRecord(b->join[l], b->side[l]);
      if (0 == code) code = (int)hipStreamWaitEvent(hs, b->join[l], 0);
    }
  }
  return (0 <= code ? lxs_hip((hipError_t)code, "batch launch") : lxs_error(code, "batch launch"));
}

LIBXSTREAM_API int libsmm_acc_batch_process(libsmm_acc_batch_t* batch, void* stream) {
  const hipStream_t hs = lxs_stream_hip((const libxstream_stream_t*)stream);
  void* token;
  int result;
  if (NULL == batch) return lxs_error(EXIT_FAILURE, "libsmm_acc_batch_process: NULL batch");
  result = lxs_ensure_device();
  if (EXIT_SUCCESS != result) return result;
  token = lxs_profile_begin(hs);
  result = lxs_batch_enqueue(batch, hs);
  lxs_profile_end(token, hs, "batch", batch->flops, batch->bytes);
  return result;
}

LIBXSTREAM_API int libsmm_acc_batch_replay(libsmm_acc_batch_t* batch, void* stream) {
  const hipStream_t hs = lxs_stream_hip((const libxstream_stream_t*)stream);
  void* token;
  int result;
  if (NULL == batch) return lxs_error(EXIT_FAILURE, "libsmm_acc_batch_replay: NULL batch");
  result = lxs_ensure_device();
  if (EXIT_SUCCESS == result && NULL == batch->exec) { /* capture the launch once */
    if (NULL == batch->capture) {
      result = lxs_hip(hipStreamCreateWithFlags(&batch->capture, hipStreamNonBlocking), "hipStreamCreate");
    }
    if (EXIT_SUCCESS == result) {
      result = lxs_hip(hipStreamBeginCapture(batch->capture, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    }
    if (EXIT_SUCCESS == result) {
      const int launched = lxs_batch_enqueue(batch, batch->capture);
      const int ended = lxs_hip(hipStreamEndCapture(batch->capture, &batch->graph), "hipStreamEndCapture");
      result = (EXIT_SUCCESS != launched ? launched : ended);
    }
    if (EXIT_SUCCESS == result) {
      result = lxs_hip(hipGraphInstantiate(&batch->exec, batch->graph, NULL, NULL, 0), "hipGraphInstantiate");
    }
  }
  if (EXIT_SUCCESS != result) return result;
  token = lxs_profile_begin(hs);
  result = lxs_hip(hipGraphLaunch(batch->exec, hs), "hipGraphLaunch");
  lxs_profile_end(token, hs, "batch", batch->flops, batch->bytes);
  return result;
}

LIBXSTREAM_API int libsmm_acc_batch_info(const libsmm_acc_batch_t* batch, int* ntasks, double* flops, double* bytes) {
  if (NULL == batch) return lxs_error(EXIT_FAILURE, "libsmm_acc_batch_info: NULL batch");
  if (NULL != ntasks) *ntasks = batch->ntasks;
  if (NULL != flops) *flops = batch->flops;
  if (NULL != bytes) *bytes = batch->bytes;
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libsmm_acc_batch_destroy(libsmm_acc_batch_t* batch) {
  int result = EXIT_SUCCESS;
  if (NULL != batch) {
    if (NULL != batch->exec) result = lxs_hip(hipGraphExecDestroy(batch->exec), "hipGraphExecDestroy");
    if (NULL != batch->graph) (void)hipGraphDestroy(batch->graph);
    if (2 == batch->per_stack) (void)hipDeviceSynchronize(); /* side streams may still run the last process */
    lxs_batch_release(batch);
  }
  return result;
}

/*
 * The batch's schedule autotuner: every candidate (chunk size, persistent work-groups per CU, and the
 * per-stack launches as one more) is planned and timed with HIP events on the caller's stacks and pools, with C redirected to scratch buffers (the
 * caller's C is neither read nor written), and the fastest is returned.
 */
LIBXSTREAM_API int libsmm_acc_batch_tune(libsmm_acc_data_t datatype, const libsmm_acc_batch_stack_t* stacks,
  int nstacks, libsmm_acc_batch_params_t* best, double* best_ms) {
  static const int scales[] = {1, 2, 4, 8, 16, 32}; /* chunk size in 23x23x23 entries x 8 */
  static const int wgs[] = {8, 16, 32};
  const int ts = (dbcsr_type_real_8 == datatype ? 8 : (dbcsr_type_real_4 == datatype ? 4 : 0));
  libsmm_acc_batch_stack_t* trial = NULL;
  void** scratch = NULL;
  hipStream_t hs = NULL;
  hipEvent_t e0 = NULL, e1 = NULL;
  double tbest = -1;
  int i, result;
  if (NULL == stacks || 0 >= nstacks || NULL == best || 0 == ts) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_batch_tune: invalid argument");
  }
  result = lxs_ensure_device();
  if (EXIT_SUCCESS != result) return result;
  trial = (libsmm_acc_batch_stack_t*)malloc(sizeof(*trial) * (size_t)nstacks);
  scratch = (void**)calloc((size_t)nstacks, sizeof(void*));
  if (NULL == trial || NULL == scratch) result = lxs_error(EXIT_FAILURE, "libsmm_acc_batch_tune: out of memory");
  for (i = 0; i < nstacks && EXIT_SUCCESS == result; ++i) { /* scratch C covering every C offset of the stack */
    const libsmm_acc_batch_stack_t* const s = stacks + i;
    long cmax = 0, j;
    trial[i] = *s;
    if (0 < s->stack_size && NULL != s->dev_param_stack) {
      int* const host = (int*)malloc(sizeof(int) * 3 * (size_t)s->stack_size);
      if (NULL == host) {
        result = lxs_error(EXIT_FAILURE, "libsmm_acc_batch_tune: out of memory");
        break;
      }
      result = lxs_hip(hipMemcpy(host, s->dev_param_stack, sizeof(int) * 3 * (size_t)s->stack_size,
        hipMemcpyDeviceToHost), "batch tune: stack download");
      for (j = 0; j < s->stack_size && EXIT_SUCCESS == result; ++j) {
        if (host[3 * j + 2] > cmax) cmax = host[3 * j + 2];
      }
      free(host);
      if (EXIT_SUCCESS == result) {
        const size_t nbytes = (size_t)ts * (size_t)(cmax - 1 + (long)s->m * s->n);
        result = lxs_hip(hipMalloc(&scratch[i], nbytes), "batch tune: scratch C");
        if (EXIT_SUCCESS == result) result = lxs_hip(hipMemset(scratch[i], 0, nbytes), "batch tune: scratch C");
        trial[i].dev_c_data = scratch[i];
      }
    }
  }
  if (EXIT_SUCCESS == result) result = lxs_hip(hipStreamCreateWithFlags(&hs, hipStreamNonBlocking), "hipStreamCreate");
  if (EXIT_SUCCESS == result) result = lxs_hip(hipEventCreate(&e0), "hipEventCreate");
  if (EXIT_SUCCESS == result) result = lxs_hip(hipEventCreate(&e1), "hipEventCreate");
  for (i = 0; i <= (int)(sizeof(scales) / sizeof(*scales)) * (int)(sizeof(wgs) / sizeof(*wgs)) + 1 && EXIT_SUCCESS == result; ++i) {
    const int last = (int)(sizeof(scales) / sizeof(*scales)) * (int)(sizeof(wgs) / sizeof(*wgs));
    libsmm_acc_batch_params_t p;
    libsmm_acc_batch_t* b = NULL;
    float ms = 0;
    int r;
    memset(&p, 0, sizeof(p));
    if (i < last) {
      p.chunk_elements = scales[i / 3] * 8 * 2 * 23 * 23;
      p.blocks_per_cu = wgs[i % 3];
    }
    else p.per_stack = 1 + (i - last); /* one launch per stack with its own configuration, in order or in lanes */
    result = libsmm_acc_batch_create_ex(datatype, trial, nstacks, &p, &b);
    if (EXIT_SUCCESS == result) result = lxs_batch_enqueue(b, hs); /* warm-up */
    if (EXIT_SUCCESS == result) result = lxs_hip(hipEventRecord(e0, hs), "hipEventRecord");
    for (r = 0; r < 3 && EXIT_SUCCESS == result; ++r) result = lxs_batch_enqueue(b, hs);
    if (EXIT_SUCCESS == result) result = lxs_hip(hipEventRecord(e1, hs), "hipEventRecord");
    if (EXIT_SUCCESS == result) result = lxs_hip(hipEventSynchronize(e1), "hipEventSynchronize");
    if (EXIT_SUCCESS == result) result = lxs_hip(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    if (NULL != b) libsmm_acc_batch_destroy(b);
    if (EXIT_SUCCESS == result && (0 > tbest || ms / 3.0 < tbest)) {
      tbest = ms / 3.0;
      *best = p;
    }
  }
  if (NULL != hs) (void)hipStreamSynchronize(hs);
  if (NULL != e0) (void)hipEventDestroy(e0);
  if (NULL != e1) (void)hipEventDestroy(e1);
  if (NULL != hs) (void)hipStreamDestroy(hs);
  for (i = 0; i < nstacks && NULL != scratch; ++i) {
    if (NULL != scratch[i]) (void)hipFree(scratch[i]);
  }
  free(scratch);
  free(trial);
  if (EXIT_SUCCESS == result && NULL != best_ms) *best_ms = tbest;
  return result;
}
