/*
 * libxstream runtime over HIP: initialization, devices, streams, events, memory, errors and
 * launch profiling. Replaces src/libxstream.c (device/launch parts), src/libxstream_stream.c,
 * src/libxstream_event.c and src/libxstream_mem.c of the reference with HIP runtime calls.
 */
#include "lxs_internal.h"
#include <math.h>
#include <string.h>

lxs_config_t lxs_cfg = {.device = -1, .lock = PTHREAD_MUTEX_INITIALIZER};

static __thread int lxs_tls_code = 0;
static __thread char lxs_tls_msg[256];
static __thread int lxs_tls_device = -1;

int lxs_error(int code, const char* what) {
  if (0 != code) {
    lxs_tls_code = code;
    snprintf(lxs_tls_msg, sizeof(lxs_tls_msg), "%s", NULL != what ? what : "");
    if (0 != lxs_cfg.verbosity) fprintf(stderr, "ERROR ACC/HIP: %s (code=%i)\n", lxs_tls_msg, code);
  }
  return code;
}

int lxs_hip(hipError_t code, const char* what) {
  if (hipSuccess == code) return EXIT_SUCCESS;
  {
    char buffer[256];
    snprintf(buffer, sizeof(buffer), "%s: %s", NULL != what ? what : "HIP", hipGetErrorString(code));
    return lxs_error((int)code, buffer);
  }
}

LIBXSTREAM_API int libxstream_error_consume(char* message, size_t size) {
  const int code = lxs_tls_code;
  if (NULL != message && 0 < size) snprintf(message, size, "%s", 0 != code ? lxs_tls_msg : "");
  lxs_tls_code = 0;
  lxs_tls_msg[0] = '\0';
  return code;
}

int lxs_ensure_device(void) {
  const int device = lxs_cfg.device;
  if (0 > device) return lxs_error(EXIT_FAILURE, "no active device");
  if (lxs_tls_device != device) {
    const int result = lxs_hip(hipSetDevice(device), "hipSetDevice");
    if (EXIT_SUCCESS != result) return result;
    lxs_tls_device = device;
  }
  return EXIT_SUCCESS;
}

hipStream_t lxs_stream_hip(const libxstream_stream_t* stream) {
  return (NULL != stream ? stream->hip : hipStreamPerThread);
}

static int lxs_getenv_int(const char* name, int dflt) {
  const char* const value = getenv(name);
  return (NULL != value && '\0' != *value) ? atoi(value) : dflt;
}

/* initialization ------------------------------------------------------------------------------ */

LIBXSTREAM_API void libxstream_init_config_default(libxstream_init_config_t* cfg) {
  if (NULL != cfg) {
    cfg->usm = -1;
    cfg->device = -1;
    cfg->verbosity = -1;
    cfg->subbuffer = -1;
  }
}

static int lxs_activate(int device_id) {
  hipDeviceProp_t props;
  int result;
  if (0 > device_id || device_id >= lxs_cfg.ndevices) return lxs_error(EXIT_FAILURE, "invalid device index");
  result = lxs_hip(hipSetDevice(device_id), "hipSetDevice");
  if (EXIT_SUCCESS == result) result = lxs_hip(hipGetDeviceProperties(&props, device_id), "hipGetDeviceProperties");
  if (EXIT_SUCCESS == result) {
    lxs_cfg.device = device_id;
    lxs_cfg.ncus = props.multiProcessorCount;
    snprintf(lxs_cfg.arch, sizeof(lxs_cfg.arch), "%s", props.gcnArchName);
    snprintf(lxs_cfg.devname, sizeof(lxs_cfg.devname), "%s %s", props.name, props.gcnArchName);
    lxs_tls_device = device_id;
    if (2 <= lxs_cfg.verbosity) {
      fprintf(stderr, "INFO ACC/HIP: ndevices=%i device%i=\"%s\" arch=%s cus=%i\n", lxs_cfg.ndevices, device_id,
        props.name, props.gcnArchName, props.multiProcessorCount);
    }
  }
  return result;
}

LIBXSTREAM_API int libxstream_init_config(const libxstream_init_config_t* cfg) {
  int result = EXIT_SUCCESS;
  pthread_mutex_lock(&lxs_cfg.lock);
  if (0 == lxs_cfg.ninit++) {
    const int env_device = lxs_getenv_int("LIBXSTREAM_DEVICE", 0);
    const int device = (NULL != cfg && 0 <= cfg->device) ? cfg->device : env_device;
    int ndevices = 0;
    lxs_cfg.verbosity = (NULL != cfg && 0 <= cfg->verbosity) ? cfg->verbosity
                                                             : lxs_getenv_int("LIBXSTREAM_VERBOSE", 0);
    lxs_cfg.profile = lxs_getenv_int("LIBXSTREAM_PROFILE", 0);
    lxs_cfg.profile_mem = lxs_getenv_int("LIBXSTREAM_PROFILE_MEM", 0);
    lxs_cfg.split = lxs_getenv_int("LIBXSTREAM_SMM_SPLIT", 2);
    lxs_cfg.split_min = lxs_getenv_int("LIBXSTREAM_SMM_SPLIT_MIN", 0);
    lxs_cfg.grouped = lxs_getenv_int("LIBXSTREAM_SMM_GROUPED", 0);
    /* no visible device is not an error here: libxstream_device_count then reports zero */
    if (hipSuccess != hipGetDeviceCount(&ndevices)) {
      (void)hipGetLastError();
      ndevices = 0;
    }
    lxs_cfg.ndevices = ndevices;
    lxs_cfg.device = -1;
    if (0 < ndevices) result = lxs_activate(device < ndevices ? device : 0);
    if (EXIT_SUCCESS != result) lxs_cfg.ninit = 0;
  }
  pthread_mutex_unlock(&lxs_cfg.lock);
  return result;
}

LIBXSTREAM_API int libxstream_init(void) { return libxstream_init_config(NULL); }

LIBXSTREAM_API int libxstream_finalize(void) {
  int result = EXIT_SUCCESS;
  pthread_mutex_lock(&lxs_cfg.lock);
  if (0 < lxs_cfg.ninit && 0 == --lxs_cfg.ninit) {
    lxs_profile_finalize();
    lxs_cfg.device = -1;
  }
  pthread_mutex_unlock(&lxs_cfg.lock);
  return result;
}

/* devices ------------------------------------------------------------------------------------- */

LIBXSTREAM_API int libxstream_device_count(int* ndevices) {
  if (NULL == ndevices) return lxs_error(EXIT_FAILURE, "libxstream_device_count: NULL argument");
  *ndevices = lxs_cfg.ndevices;
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libxstream_device_set_active(int device_id) {
  int result;
  pthread_mutex_lock(&lxs_cfg.lock);
  result = (lxs_cfg.device == device_id && lxs_tls_device == device_id) ? EXIT_SUCCESS : lxs_activate(device_id);
  pthread_mutex_unlock(&lxs_cfg.lock);
  return result;
}

LIBXSTREAM_API int libxstream_device_sync(void) {
  int result = lxs_ensure_device();
  if (EXIT_SUCCESS == result) result = lxs_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return result;
}

LIBXSTREAM_API int libxstream_device_info(int* ncus, char* name, size_t name_size) {
  if (0 > lxs_cfg.device) return lxs_error(EXIT_FAILURE, "no active device");
  if (NULL != ncus) *ncus = lxs_cfg.ncus;
  if (NULL != name && 0 < name_size) snprintf(name, name_size, "%s", lxs_cfg.arch);
  return EXIT_SUCCESS;
}

/* streams ------------------------------------------------------------------------------------- */

LIBXSTREAM_API int libxstream_stream_create(libxstream_stream_t** stream_p, const char* name, int flags) {
  libxstream_stream_t* stream;
  int result;
  if (NULL == stream_p) return lxs_error(EXIT_FAILURE, "libxstream_stream_create: NULL argument");
  *stream_p = NULL;
  result = lxs_ensure_device();
  if (EXIT_SUCCESS != result) return result;
  stream = (libxstream_stream_t*)calloc(1, sizeof(*stream));
  if (NULL == stream) return lxs_error(EXIT_FAILURE, "libxstream_stream_create: out of memory");
  /* default priority for every stream; non-blocking w.r.t. the legacy null stream like an OpenCL queue */
  result = lxs_hip(hipStreamCreateWithFlags(&stream->hip, hipStreamNonBlocking), "hipStreamCreate");
  if (EXIT_SUCCESS == result) {
    stream->device = lxs_cfg.device;
    stream->flags = flags;
    snprintf(stream->name, sizeof(stream->name), "%s", NULL != name ? name : "");
    *stream_p = stream;
  }
  else {
    free(stream);
  }
  return result;
}

LIBXSTREAM_API int libxstream_stream_destroy(libxstream_stream_t* stream) {
  int result = EXIT_SUCCESS;
  if (NULL != stream) {
    result = lxs_hip(hipStreamDestroy(stream->hip), "hipStreamDestroy");
    free(stream);
  }
  return result;
}

LIBXSTREAM_API int libxstream_stream_sync(libxstream_stream_t* stream) {
  int result = lxs_ensure_device();
  if (EXIT_SUCCESS == result) result = lxs_hip(hipStreamSynchronize(lxs_stream_hip(stream)), "hipStreamSynchronize");
  return result;
}

LIBXSTREAM_API int libxstream_stream_wait_event(libxstream_stream_t* stream, libxstream_event_t* event) {
  int result;
  if (NULL == event) return lxs_error(EXIT_FAILURE, "libxstream_stream_wait_event: NULL event");
  result = lxs_ensure_device();
  if (EXIT_SUCCESS == result) {
    result = lxs_hip(hipStreamWaitEvent(lxs_stream_hip(stream), event->hip, 0), "hipStreamWaitEvent");
  }
  return result;
}

LIBXSTREAM_API int libxstream_stream_set_profiling(libxstream_stream_t* stream) {
  (void)stream;
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libxstream_stream_priority_range(int* least, int* greatest) {
  if (NULL == least && NULL == greatest) return lxs_error(EXIT_FAILURE, "libxstream_stream_priority_range");
  if (NULL != least) *least = -1;
  if (NULL != greatest) *greatest = -1;
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libxstream_stream_native(libxstream_stream_t* stream, void** hip_stream) {
  if (NULL == hip_stream) return lxs_error(EXIT_FAILURE, "libxstream_stream_native: NULL argument");
  *hip_stream = (void*)lxs_stream_hip(stream);
  return EXIT_SUCCESS;
}

/* events -------------------------------------------------------------------------------------- */

LIBXSTREAM_API int libxstream_event_create(libxstream_event_t** event_p) {
  libxstream_event_t* event;
  int result;
  if (NULL == event_p) return lxs_error(EXIT_FAILURE, "libxstream_event_create: NULL argument");
  *event_p = NULL;
  result = lxs_ensure_device();
  if (EXIT_SUCCESS != result) return result;
  event = (libxstream_event_t*)calloc(1, sizeof(*event));
  if (NULL == event) return lxs_error(EXIT_FAILURE, "libxstream_event_create: out of memory");
  result = lxs_hip(hipEventCreateWithFlags(&event->hip, hipEventDefault), "hipEventCreate");
  if (EXIT_SUCCESS == result) {
    *event_p = event;
  }
  else {
    free(event);
  }
  return result;
}

LIBXSTREAM_API int libxstream_event_destroy(libxstream_event_t* event) {
  int result = EXIT_SUCCESS;
  if (NULL != event) {
    result = lxs_hip(hipEventDestroy(event->hip), "hipEventDestroy");
    free(event);
  }
  return result;
}

LIBXSTREAM_API int libxstream_event_record(libxstream_event_t* event, libxstream_stream_t* stream) {
  int result;
  if (NULL == event) return lxs_error(EXIT_FAILURE, "libxstream_event_record: NULL event");
  result = lxs_ensure_device();
  if (EXIT_SUCCESS == result) result = lxs_hip(hipEventRecord(event->hip, lxs_stream_hip(stream)), "hipEventRecord");
  return result;
}

LIBXSTREAM_API int libxstream_event_query(libxstream_event_t* event, libxstream_bool_t* has_occurred) {
  hipError_t code;
  if (NULL == event || NULL == has_occurred) return lxs_error(EXIT_FAILURE, "libxstream_event_query: NULL argument");
  code = hipEventQuery(event->hip);
  if (hipSuccess == code) {
    *has_occurred = 1;
    return EXIT_SUCCESS;
  }
  if (hipErrorNotReady == code) {
    *has_occurred = 0;
    return EXIT_SUCCESS;
  }
  *has_occurred = 0;
  return lxs_hip(code, "hipEventQuery");
}

LIBXSTREAM_API int libxstream_event_sync(libxstream_event_t* event) {
  if (NULL == event) return lxs_error(EXIT_FAILURE, "libxstream_event_sync: NULL event");
  return lxs_hip(hipEventSynchronize(event->hip), "hipEventSynchronize");
}

LIBXSTREAM_API int libxstream_event_elapsed(libxstream_event_t* begin, libxstream_event_t* end, double* ms) {
  float value = 0;
  int result;
  if (NULL == begin || NULL == end || NULL == ms) return lxs_error(EXIT_FAILURE, "libxstream_event_elapsed");
  result = lxs_hip(hipEventElapsedTime(&value, begin->hip, end->hip), "hipEventElapsedTime");
  *ms = (EXIT_SUCCESS == result ? (double)value : 0.0);
  return result;
}

/* memory -------------------------------------------------------------------------------------- */

LIBXSTREAM_API int libxstream_mem_allocate(void** dev_mem, size_t nbytes) {
  int result;
  if (NULL == dev_mem) return lxs_error(EXIT_FAILURE, "libxstream_mem_allocate: NULL argument");
  *dev_mem = NULL;
  if (0 == nbytes) return EXIT_SUCCESS;
  result = lxs_ensure_device();
  if (EXIT_SUCCESS == result) result = lxs_hip(hipMalloc(dev_mem, nbytes), "hipMalloc");
  if (EXIT_SUCCESS != result) *dev_mem = NULL;
  return result;
}

LIBXSTREAM_API int libxstream_mem_deallocate(void* dev_mem) {
  int result = EXIT_SUCCESS;
  if (NULL != dev_mem) {
    result = lxs_ensure_device();
    if (EXIT_SUCCESS == result) result = lxs_hip(hipFree(dev_mem), "hipFree");
  }
  return result;
}

LIBXSTREAM_API int libxstream_mem_dev_allocate_hint(void** dev_mem, size_t nbytes, libxstream_mem_hint_t hint) {
  (void)hint;
  return libxstream_mem_allocate(dev_mem, nbytes);
}

LIBXSTREAM_API int libxstream_mem_dev_deallocate_hint(void* dev_mem) { return libxstream_mem_deallocate(dev_mem); }

LIBXSTREAM_API int libxstream_mem_offset(void** dev_mem, void* other, size_t lb) {
  if (NULL == dev_mem) return lxs_error(EXIT_FAILURE, "libxstream_mem_offset: NULL argument");
  if (NULL != other || 0 == lb) {
    *dev_mem = (char*)other + lb;
    return EXIT_SUCCESS;
  }
  *dev_mem = NULL;
  return lxs_error(EXIT_FAILURE, "libxstream_mem_offset: offset into NULL");
}

LIBXSTREAM_API int libxstream_mem_info(size_t* mem_free, size_t* mem_total) {
  size_t f = 0, t = 0;
  int result = lxs_ensure_device();
  if (EXIT_SUCCESS == result) result = lxs_hip(hipMemGetInfo(&f, &t), "hipMemGetInfo");
  if (NULL != mem_free) *mem_free = f;
  if (NULL != mem_total) *mem_total = t;
  return result;
}

LIBXSTREAM_API int libxstream_mem_host_allocate(void** host_mem, size_t nbytes, libxstream_stream_t* stream) {
  int result;
  (void)stream;
  if (NULL == host_mem) return lxs_error(EXIT_FAILURE, "libxstream_mem_host_allocate: NULL argument");
  *host_mem = NULL;
  if (0 == nbytes) return EXIT_SUCCESS;
  result = lxs_ensure_device();
  if (EXIT_SUCCESS == result) result = lxs_hip(hipHostMalloc(host_mem, nbytes, hipHostMallocDefault), "hipHostMalloc");
  if (EXIT_SUCCESS != result) *host_mem = NULL;
  return result;
}

LIBXSTREAM_API int libxstream_mem_host_deallocate(void* host_mem, libxstream_stream_t* stream) {
  (void)stream;
  return (NULL != host_mem ? lxs_hip(hipHostFree(host_mem), "hipHostFree") : EXIT_SUCCESS);
}

static int lxs_copy(void* dst, const void* src, size_t nbytes, hipMemcpyKind kind, libxstream_stream_t* stream,
  const char* what) {
  const hipStream_t hs = lxs_stream_hip(stream);
  int result;
  if (0 == nbytes) return EXIT_SUCCESS;
  if (NULL == dst || NULL == src) return lxs_error(EXIT_FAILURE, what);
  result = lxs_ensure_device();
  if (EXIT_SUCCESS == result) {
    void* const token = lxs_profile_begin_mem(hs);
    result = lxs_hip(hipMemcpyAsync(dst, src, nbytes, kind, hs), what);
    lxs_profile_end(token, hs,
      hipMemcpyHostToDevice == kind ? "H2D" : (hipMemcpyDeviceToHost == kind ? "D2H" : "D2D"), 0, (double)nbytes);
  }
  /* a NULL stream makes the copy blocking (src/libxstream_mem.c:1105) */
  if (EXIT_SUCCESS == result && NULL == stream) result = lxs_hip(hipStreamSynchronize(hs), what);
  return result;
}

LIBXSTREAM_API int libxstream_mem_copy_h2d(const void* host_mem, void* dev_mem, size_t nbytes,
  libxstream_stream_t* stream) {
  return lxs_copy(dev_mem, host_mem, nbytes, hipMemcpyHostToDevice, stream, "copy h2d");
}

LIBXSTREAM_API int libxstream_mem_copy_d2h(const void* dev_mem, void* host_mem, size_t nbytes,
  libxstream_stream_t* stream) {
  return lxs_copy(host_mem, dev_mem, nbytes, hipMemcpyDeviceToHost, stream, "copy d2h");
}

LIBXSTREAM_API int libxstream_mem_copy_d2d(const void* devmem_src, void* devmem_dst, size_t nbytes,
  libxstream_stream_t* stream) {
  return lxs_copy(devmem_dst, devmem_src, nbytes, hipMemcpyDeviceToDevice, stream, "copy d2d");
}

LIBXSTREAM_API int libxstream_mem_zero(void* dev_mem, size_t offset, size_t nbytes, libxstream_stream_t* stream) {
  const hipStream_t hs = lxs_stream_hip(stream);
  int result;
  if (0 == nbytes) return EXIT_SUCCESS;
  if (NULL == dev_mem) return lxs_error(EXIT_FAILURE, "libxstream_mem_zero: NULL pointer");
  result = lxs_ensure_device();
  if (EXIT_SUCCESS == result) {
    void* const token = lxs_profile_begin_mem(hs);
    const int code = lxs_zero_launch((char*)dev_mem + offset, nbytes, (void*)hs);
    lxs_profile_end(token, hs, "ZERO", 0, (double)nbytes);
    result = (0 == code ? EXIT_SUCCESS : lxs_error(code, "zero fill launch"));
  }
  if (EXIT_SUCCESS == result && NULL == stream) result = lxs_hip(hipStreamSynchronize(hs), "zero fill");
  return result;
}

/* launch profiling (LIBXSTREAM_PROFILE, src/libxstream.c:853-1015, :2629-2781) and transfer profiling
   (LIBXSTREAM_PROFILE_MEM, src/libxstream_mem.c:805-893) ------------------------------------------------
   Every profiled launch or copy is bracketed by two HIP events on its stream. A completed record becomes a
   sample {duration, flops, bytes} of its name; durations shorter than LXS_PROFILE_TICKS ticks of the device's
   wall clock are counted as discarded (quantisation, not a measurement). Each name keeps its totals and the
   latest LXS_PROFILE_SAMPLES samples, from which the median and a histogram are formed at print time; rates
   are always taken from a sample's own work and duration. */

#define LXS_PROFILE_NAMES 24
#define LXS_PROFILE_SAMPLES 65536
#define LXS_PROFILE_TICKS 10
typedef struct lxs_record_t {
  hipEvent_t begin, end;
  int name;
  double flops, bytes;
} lxs_record_t;
typedef struct lxs_sample_t {
  float ms;
  double flops, bytes;
} lxs_sample_t;
typedef struct lxs_stat_t {
  char name[32];
  long count, discarded;
  double ms, flops, bytes, min_ms, max_ms;
  lxs_sample_t* samples; /* ring of the latest LXS_PROFILE_SAMPLES samples */
  int transfer;          /* a copy or fill: samples are keyed by size */
} lxs_stat_t;

static pthread_mutex_t lxs_prof_lock = PTHREAD_MUTEX_INITIALIZER;
static lxs_record_t* lxs_prof_pending = NULL;
static int lxs_prof_npending = 0, lxs_prof_cap = 0;
static hipEvent_t* lxs_prof_pool = NULL;
static int lxs_prof_npool = 0, lxs_prof_poolcap = 0;
static lxs_stat_t lxs_prof_stats[LXS_PROFILE_NAMES];
static int lxs_prof_nstats = 0;
static double lxs_prof_timer_ns = 0; /* tick of the event clock (0: unknown) */

static hipEvent_t lxs_prof_event(void) { /* lock held */
  hipEvent_t event = NULL;
  if (0 < lxs_prof_npool) return lxs_prof_pool[--lxs_prof_npool];
  if (hipSuccess != hipEventCreateWithFlags(&event, hipEventDefault)) event = NULL;
  return event;
}

static void lxs_prof_release(hipEvent_t event) { /* lock held */
  if (lxs_prof_npool == lxs_prof_poolcap) {
    const int cap = (0 < lxs_prof_poolcap ? 2 * lxs_prof_poolcap : 64);
    hipEvent_t* const pool = (hipEvent_t*)realloc(lxs_prof_pool, sizeof(hipEvent_t) * (size_t)cap);
    if (NULL == pool) {
      (void)hipEventDestroy(event);
      return;
    }
    lxs_prof_pool = pool;
    lxs_prof_poolcap = cap;
  }
  lxs_prof_pool[lxs_prof_npool++] = event;
}

static int lxs_prof_is_transfer(const char* name) {
  return 0 == strcmp(name, "H2D") || 0 == strcmp(name, "D2H") || 0 == strcmp(name, "D2D") || 0 == strcmp(name, "ZERO");
}

static int lxs_prof_name(const char* kernel) { /* lock held */
  int i;
  for (i = 0; i < lxs_prof_nstats; ++i) {
    if (0 == strcmp(lxs_prof_stats[i].name, kernel)) return i;
  }
  if (LXS_PROFILE_NAMES <= lxs_prof_nstats) return LXS_PROFILE_NAMES - 1;
  free(lxs_prof_stats[i].samples);
  memset(lxs_prof_stats + i, 0, sizeof(*lxs_prof_stats));
  snprintf(lxs_prof_stats[i].name, sizeof(lxs_prof_stats[i].name), "%s", kernel);
  lxs_prof_stats[i].transfer = lxs_prof_is_transfer(kernel);
  return lxs_prof_nstats++;
}

static double lxs_prof_floor_ms(void) {
  if (0 == lxs_prof_timer_ns) {
    int khz = 0;
    /* the events' clock: the device's constant-rate wall clock */
    if (0 <= lxs_cfg.device && hipSuccess == hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, lxs_cfg.device) &&
        0 < khz) {
      lxs_prof_timer_ns = 1e6 / khz;
    }
    else {
      (void)hipGetLastError();
      lxs_prof_timer_ns = 10.0;
    }
  }
  return 1e-6 * LXS_PROFILE_TICKS * lxs_prof_timer_ns;
}

static void lxs_prof_push(lxs_stat_t* st, double ms, double flops, double bytes) { /* lock held */
  if (ms < lxs_prof_floor_ms()) {
    ++st->discarded;
    if (0 > lxs_cfg.profile) fprintf(stderr, "PROF ACC/HIP: %s ms=%.4f (below the floor, discarded)\n", st->name, ms);
    return;
  }
  if (NULL == st->samples) st->samples = (lxs_sample_t*)malloc(sizeof(lxs_sample_t) * LXS_PROFILE_SAMPLES);
  if (NULL != st->samples) {
    lxs_sample_t* const x = st->samples + (st->count % LXS_PROFILE_SAMPLES);
    x->ms = (float)ms;
    x->flops = flops;
    x->bytes = bytes;
  }
  st->min_ms = (0 == st->count || ms < st->min_ms ? ms : st->min_ms);
  st->max_ms = (0 == st->count || ms > st->max_ms ? ms : st->max_ms);
  ++st->count;
  st->ms += ms;
  st->flops += flops;
  st->bytes += bytes;
  if (0 > lxs_cfg.profile) fprintf(stderr, "PROF ACC/HIP: %s ms=%.4f\n", st->name, ms);
}

/* fold completed records into the statistics; blocking waits for all of them (lock held) */
static void lxs_prof_drain(int blocking) {
  int i, j = 0;
  for (i = 0; i < lxs_prof_npending; ++i) {
    lxs_record_t* const r = lxs_prof_pending + i;
    const hipError_t state = (0 != blocking ? hipEventSynchronize(r->end) : hipEventQuery(r->end));
    if (hipSuccess == state) {
      float ms = 0;
      if (hipSuccess == hipEventElapsedTime(&ms, r->begin, r->end)) {
        lxs_prof_push(lxs_prof_stats + r->name, ms, r->flops, r->bytes);
      }
      lxs_prof_release(r->begin);
      lxs_prof_release(r->end);
    }
    else {
      (void)hipGetLastError();
      lxs_prof_pending[j++] = *r;
    }
  }
  lxs_prof_npending = j;
}

static int lxs_prof_cmp_ms(const void* a, const void* b) {
  const float x = ((const lxs_sample_t*)a)->ms, y = ((const lxs_sample_t*)b)->ms;
  return (x < y ? -1 : (x > y ? 1 : 0));
}

static int lxs_prof_cmp_bytes(const void* a, const void* b) {
  const double x = ((const lxs_sample_t*)a)->bytes, y = ((const lxs_sample_t*)b)->bytes;
  return (x < y ? -1 : (x > y ? 1 : 0));
}

/* the kept samples of a statistic, sorted by duration (kernels) or by size (transfers); caller frees */
static lxs_sample_t* lxs_prof_sorted(const lxs_stat_t* st, long* n) { /* lock held */
  const long count = (st->count < LXS_PROFILE_SAMPLES ? st->count : LXS_PROFILE_SAMPLES);
  lxs_sample_t* const x = (0 < count && NULL != st->samples ? (lxs_sample_t*)malloc(sizeof(lxs_sample_t) * count) : NULL);
  *n = 0;
  if (NULL == x) return NULL;
  memcpy(x, st->samples, sizeof(lxs_sample_t) * (size_t)count);
  qsort(x, (size_t)count, sizeof(lxs_sample_t), 0 != st->transfer ? lxs_prof_cmp_bytes : lxs_prof_cmp_ms);
  *n = count;
  return x;
}

static void* lxs_profile_begin_any(hipStream_t stream) {
  hipEvent_t event;
  pthread_mutex_lock(&lxs_prof_lock);
  event = lxs_prof_event();
  pthread_mutex_unlock(&lxs_prof_lock);
  if (NULL != event && hipSuccess != hipEventRecord(event, stream)) {
    pthread_mutex_lock(&lxs_prof_lock);
    lxs_prof_release(event);
    pthread_mutex_unlock(&lxs_prof_lock);
    event = NULL;
  }
  return (void*)event;
}

void* lxs_profile_begin(hipStream_t stream) {
  if (0 == lxs_cfg.profile) return NULL;
  return lxs_profile_begin_any(stream);
}

void* lxs_profile_begin_mem(hipStream_t stream) {
  if (0 == lxs_cfg.profile_mem) return NULL;
  return lxs_profile_begin_any(stream);
}

void lxs_profile_end(void* token, hipStream_t stream, const char* kernel, double flops, double bytes) {
  hipEvent_t end;
  if (NULL == token) return;
  pthread_mutex_lock(&lxs_prof_lock);
  end = lxs_prof_event();
  if (NULL != end && hipSuccess == hipEventRecord(end, stream)) {
    if (lxs_prof_npending == lxs_prof_cap) {
      const int cap = (0 < lxs_prof_cap ? 2 * lxs_prof_cap : 256);
      lxs_record_t* const p = (lxs_record_t*)realloc(lxs_prof_pending, sizeof(lxs_record_t) * (size_t)cap);
      if (NULL != p) {
        lxs_prof_pending = p;
        lxs_prof_cap = cap;
      }
    }
    if (lxs_prof_npending < lxs_prof_cap) {
      lxs_record_t* const r = lxs_prof_pending + lxs_prof_npending++;
      r->begin = (hipEvent_t)token;
      r->end = end;
      r->name = lxs_prof_name(kernel);
      r->flops = flops;
      r->bytes = bytes;
      token = NULL;
      end = NULL;
    }
    if (1024 <= lxs_prof_npending) lxs_prof_drain(0);
  }
  if (NULL != token) lxs_prof_release((hipEvent_t)token);
  if (NULL != end) lxs_prof_release(end);
  pthread_mutex_unlock(&lxs_prof_lock);
}

LIBXSTREAM_API int libxstream_profile_enable(int enable) {
  lxs_cfg.profile = enable;
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libxstream_profile_mem_enable(int enable) {
  lxs_cfg.profile_mem = enable;
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libxstream_profile_reset(void) {
  int i;
  pthread_mutex_lock(&lxs_prof_lock);
  lxs_prof_drain(1);
  for (i = 0; i < lxs_prof_nstats; ++i) {
    free(lxs_prof_stats[i].samples);
    lxs_prof_stats[i].samples = NULL;
  }
  lxs_prof_nstats = 0;
  pthread_mutex_unlock(&lxs_prof_lock);
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libxstream_profile_get(const char* kernel, long* count, double* total_ms, double* total_flops,
  double* total_bytes) {
  int i, found = 0;
  if (NULL == kernel) return lxs_error(EXIT_FAILURE, "libxstream_profile_get: NULL kernel");
  pthread_mutex_lock(&lxs_prof_lock);
  lxs_prof_drain(1);
  for (i = 0; i < lxs_prof_nstats; ++i) {
    if (0 == strcmp(lxs_prof_stats[i].name, kernel)) {
      if (NULL != count) *count = lxs_prof_stats[i].count;
      if (NULL != total_ms) *total_ms = lxs_prof_stats[i].ms;
      if (NULL != total_flops) *total_flops = lxs_prof_stats[i].flops;
      if (NULL != total_bytes) *total_bytes = lxs_prof_stats[i].bytes;
      found = 1;
    }
  }
  pthread_mutex_unlock(&lxs_prof_lock);
  if (0 == found) {
    if (NULL != count) *count = 0;
    if (NULL != total_ms) *total_ms = 0;
    if (NULL != total_flops) *total_flops = 0;
    if (NULL != total_bytes) *total_bytes = 0;
  }
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libxstream_profile_stats(const char* name, double* min_ms, double* median_ms, double* max_ms,
  double* median_rate, long* discarded) {
  int i, found = -1;
  if (NULL == name) return lxs_error(EXIT_FAILURE, "libxstream_profile_stats: NULL name");
  pthread_mutex_lock(&lxs_prof_lock);
  lxs_prof_drain(1);
  for (i = 0; i < lxs_prof_nstats && 0 > found; ++i) {
    if (0 == strcmp(lxs_prof_stats[i].name, name)) found = i;
  }
  if (NULL != min_ms) *min_ms = 0;
  if (NULL != median_ms) *median_ms = 0;
  if (NULL != max_ms) *max_ms = 0;
  if (NULL != median_rate) *median_rate = 0;
  if (NULL != discarded) *discarded = 0;
  if (0 <= found) {
    const lxs_stat_t* const st = lxs_prof_stats + found;
    long n = 0;
    lxs_sample_t* const x = lxs_prof_sorted(st, &n);
    if (NULL != min_ms) *min_ms = st->min_ms;
    if (NULL != max_ms) *max_ms = st->max_ms;
    if (NULL != discarded) *discarded = st->discarded;
    if (NULL != x && 0 < n) {
      const lxs_sample_t* const mid = x + n / 2; /* the median sample: by duration, or by size for transfers */
      if (NULL != median_ms) *median_ms = mid->ms;
      if (NULL != median_rate && 0 < mid->ms) { /* GB/s of a transfer, GFLOP/s of a kernel */
        *median_rate = 1e-6 * (0 != st->transfer ? mid->bytes : mid->flops) / mid->ms;
      }
    }
    free(x);
  }
  pthread_mutex_unlock(&lxs_prof_lock);
  return (0 <= found ? EXIT_SUCCESS : EXIT_FAILURE);
}

/* one statistic as the reference prints it: the median sample's duration and rates (kernels) or the rate of
   the most populated size bucket (transfers), then nbuckets rows of counts and per-bucket averages */
static void lxs_prof_print_stat(FILE* stream, const lxs_stat_t* st, int nbuckets) { /* lock held */
  long n = 0, b, i0 = 0;
  lxs_sample_t* const x = lxs_prof_sorted(st, &n);
  if (NULL == x || 0 == n) {
    free(x);
    return;
  }
  if (0 != st->transfer) { /* buckets of sizes on a log2 axis; the mode bucket gives the headline rate */
    const double lo = log2(x[0].bytes > 1 ? x[0].bytes : 1), hi = log2(x[n - 1].bytes > 1 ? x[n - 1].bytes : 1);
    double best_rate = 0;
    long best = -1;
    for (b = 0; b < nbuckets; ++b) {
      double sb = 0, sms = 0;
      long cnt = 0;
      const double edge = (b + 1 == nbuckets ? hi + 1 : lo + (hi - lo) * (b + 1) / nbuckets);
      while (i0 < n && log2(x[i0].bytes > 1 ? x[i0].bytes : 1) < edge) {
        sb += x[i0].bytes;
        sms += x[i0].ms;
        ++cnt;
        ++i0;
      }
      if (cnt > best && 0 < sms) {
        best = cnt;
        best_rate = 1e-6 * sb / sms;
      }
    }
    fprintf(stream, "PROF ACC/HIP: %s=%.1f GB/s n=%li\n", st->name, best_rate, st->count);
    for (b = 0, i0 = 0; b < nbuckets; ++b) {
      double sb = 0, sms = 0;
      long cnt = 0;
      const double edge = (b + 1 == nbuckets ? hi + 1 : lo + (hi - lo) * (b + 1) / nbuckets);
      while (i0 < n && log2(x[i0].bytes > 1 ? x[i0].bytes : 1) < edge) {
        sb += x[i0].bytes;
        sms += x[i0].ms;
        ++cnt;
        ++i0;
      }
      if (0 < cnt) {
        fprintf(stream, "  %.1f MB x%li: %.1f us %.1f GB/s\n", 1e-6 * sb / cnt, cnt, 1e3 * sms / cnt,
          0 < sms ? 1e-6 * sb / sms : 0.0);
      }
    }
  }
  else {
    const lxs_sample_t* const mid = x + n / 2;
    const double lo = x[0].ms, hi = x[n - 1].ms;
    fprintf(stream, "PROF ACC/HIP: %s=%.4f ms", st->name, (double)mid->ms);
    if (0 < mid->flops) fprintf(stream, " %.1f GFLOPS/s", 1e-6 * mid->flops / mid->ms);
    if (0 < mid->bytes) fprintf(stream, " %.1f GB/s", 1e-6 * mid->bytes / mid->ms);
    fprintf(stream, " n=%li min=%.4f max=%.4f ms\n", st->count, st->min_ms, st->max_ms);
    for (b = 0; b < nbuckets && lo < hi; ++b) {
      double sf = 0, sb = 0, sms = 0;
      long cnt = 0;
      const double edge = (b + 1 == nbuckets ? hi * 2 : lo + (hi - lo) * (b + 1) / nbuckets);
      while (i0 < n && x[i0].ms < edge) {
        sf += x[i0].flops;
        sb += x[i0].bytes;
        sms += x[i0].ms;
        ++cnt;
        ++i0;
      }
      if (0 < cnt) {
        fprintf(stream, "  <%.4f ms x%li: %.4f ms", (b + 1 == nbuckets ? hi : edge), cnt, sms / cnt);
        if (0 < sf) fprintf(stream, " %.1f GFLOPS/s", 1e-6 * sf / sms);
        if (0 < sb) fprintf(stream, " %.1f GB/s", 1e-6 * sb / sms);
        fprintf(stream, "\n");
      }
    }
  }
  free(x);
}

LIBXSTREAM_API int libxstream_profile_print(FILE* stream) {
  const int setting = (0 <= lxs_cfg.profile ? lxs_cfg.profile : -lxs_cfg.profile);
  const int nbuckets = (2 < setting ? setting : 2) + 1;
  long discarded = 0;
  int i;
  if (NULL == stream) stream = stderr;
  pthread_mutex_lock(&lxs_prof_lock);
  lxs_prof_drain(1);
  for (i = 0; i < lxs_prof_nstats; ++i) {
    lxs_prof_print_stat(stream, lxs_prof_stats + i, nbuckets);
    discarded += lxs_prof_stats[i].discarded;
  }
  if (0 < discarded) {
    fprintf(stream, "PROF ACC/HIP: discarded=%li timer=%.0fns floor=%.0fns\n", discarded, lxs_prof_timer_ns,
      LXS_PROFILE_TICKS * lxs_prof_timer_ns);
  }
  else if (0 == lxs_prof_nstats) {
    fprintf(stream, "PROF ACC/HIP: no samples recorded\n");
  }
  pthread_mutex_unlock(&lxs_prof_lock);
  return EXIT_SUCCESS;
}

void lxs_profile_finalize(void) {
  int i;
  if ((0 != lxs_cfg.profile || 0 != lxs_cfg.profile_mem) && 0 < lxs_prof_nstats) libxstream_profile_print(stderr);
  pthread_mutex_lock(&lxs_prof_lock);
  lxs_prof_drain(1);
  for (i = 0; i < lxs_prof_npool; ++i) (void)hipEventDestroy(lxs_prof_pool[i]);
  lxs_prof_npool = 0;
  for (i = 0; i < lxs_prof_nstats; ++i) {
    free(lxs_prof_stats[i].samples);
    lxs_prof_stats[i].samples = NULL;
  }
  lxs_prof_nstats = 0;
  pthread_mutex_unlock(&lxs_prof_lock);
}
