/*
 * libxstream-amd: public stream/event/memory API over HIP (MI355X, gfx950).
 *
 * Drop-in for the reference's public API, libxstream/libxstream.h:15-79 (hfp/libxstream).
 * Names, argument meaning and return conventions are kept; the implementation is new and
 * runs on the HIP runtime instead of OpenCL queues/events/cl_mem.
 *
 * Return convention (libxstream/libxstream_opencl.h:156-168 for the reference's CL_RETURN):
 *   EXIT_SUCCESS (0) on success, non-zero on failure. A HIP failure returns the positive
 *   hipError_t value; a failed argument check returns EXIT_FAILURE. The message of the most
 *   recent failure on the calling thread can be consumed with libxstream_error_consume().
 */
#ifndef LIBXSTREAM_H
#define LIBXSTREAM_H

#include <stddef.h>

#if defined(__cplusplus)
#  define LIBXSTREAM_EXTERN_C extern "C"
#else
#  define LIBXSTREAM_EXTERN_C
#endif
#define LIBXSTREAM_API LIBXSTREAM_EXTERN_C __attribute__((visibility("default")))

/** types (libxstream/libxstream.h:16-18) */
typedef int libxstream_bool_t;
typedef struct libxstream_stream_t libxstream_stream_t;
typedef struct libxstream_event_t libxstream_event_t;

/** initialization config (libxstream/libxstream.h:21-26); usm/subbuffer have no effect on HIP */
typedef struct libxstream_init_config_t {
  int usm;       /* -1: env/default; ignored (HIP device pointers are plain pointers) */
  int device;    /* -1: env LIBXSTREAM_DEVICE or first device, >=0: device index */
  int verbosity; /* -1: env LIBXSTREAM_VERBOSE or 0, >=0: explicit level */
  int subbuffer; /* -1: env/default; ignored */
} libxstream_init_config_t;

/** libxstream/libxstream.h:29 */
LIBXSTREAM_API void libxstream_init_config_default(libxstream_init_config_t* cfg);
/** libxstream/libxstream.h:31; explicit > env > default precedence as src/libxstream.c:178-189 */
LIBXSTREAM_API int libxstream_init_config(const libxstream_init_config_t* cfg);
/** libxstream/libxstream.h:33; reference-counted, repeated calls succeed (samples/smm/smm_acc.c:54-62) */
LIBXSTREAM_API int libxstream_init(void);
/** libxstream/libxstream.h:34 */
LIBXSTREAM_API int libxstream_finalize(void);

/** devices (libxstream/libxstream.h:37-39) */
LIBXSTREAM_API int libxstream_device_count(int* ndevices);
LIBXSTREAM_API int libxstream_device_set_active(int device_id);
LIBXSTREAM_API int libxstream_device_sync(void);

/** streams (libxstream/libxstream.h:42-54) */
typedef enum libxstream_stream_flags_t {
  LIBXSTREAM_STREAM_DEFAULT = 0,
  LIBXSTREAM_STREAM_LOW = 1,
  LIBXSTREAM_STREAM_HIGH = 2,
  LIBXSTREAM_STREAM_PROFILING = 4
} libxstream_stream_flags_t;
/** src/libxstream_stream.c:54-169. Every stream is created at the default priority: LOW/HIGH are
 *  accepted and recorded but not applied (streams of mixed priority are not used on MI355X). */
LIBXSTREAM_API int libxstream_stream_create(libxstream_stream_t** stream_p, const char* name, int flags);
/** src/libxstream_stream.c:172-194; NULL is accepted (no-op) */
LIBXSTREAM_API int libxstream_stream_destroy(libxstream_stream_t* stream);
/** src/libxstream_stream.c:227-248; NULL selects the calling thread's default stream */
LIBXSTREAM_API int libxstream_stream_sync(libxstream_stream_t* stream);
/** src/libxstream_event.c:46-80 */
LIBXSTREAM_API int libxstream_stream_wait_event(libxstream_stream_t* stream, libxstream_event_t* event);
/** libxstream/libxstream.h:54; all streams record timing-capable events, so this is a no-op */
LIBXSTREAM_API int libxstream_stream_set_profiling(libxstream_stream_t* stream);
/** src/libxstream_stream.c:197-224; reports least == greatest (no priority range in use) */
LIBXSTREAM_API int libxstream_stream_priority_range(int* least, int* greatest);

/** events (libxstream/libxstream.h:57-61; src/libxstream_event.c:14-169) */
LIBXSTREAM_API int libxstream_event_create(libxstream_event_t** event_p);
LIBXSTREAM_API int libxstream_event_destroy(libxstream_event_t* event);
LIBXSTREAM_API int libxstream_event_record(libxstream_event_t* event, libxstream_stream_t* stream);
LIBXSTREAM_API int libxstream_event_query(libxstream_event_t* event, libxstream_bool_t* has_occurred);
LIBXSTREAM_API int libxstream_event_sync(libxstream_event_t* event);

/** memory (libxstream/libxstream.h:64-79; src/libxstream_mem.c) */
/** src/libxstream_mem.c:896-946: nbytes == 0 yields NULL and success */
LIBXSTREAM_API int libxstream_mem_allocate(void** dev_mem, size_t nbytes);
LIBXSTREAM_API int libxstream_mem_deallocate(void* dev_mem);
/** src/libxstream_mem.c:987-999: *dev_mem = (char*)other + lb; NULL other with lb != 0 fails */
LIBXSTREAM_API int libxstream_mem_offset(void** dev_mem, void* other, size_t lb);
LIBXSTREAM_API int libxstream_mem_info(size_t* mem_free, size_t* mem_total);
/** src/libxstream_mem.c:702-772: pinned host memory (hipHostMalloc) */
LIBXSTREAM_API int libxstream_mem_host_allocate(void** host_mem, size_t nbytes, libxstream_stream_t* stream);
LIBXSTREAM_API int libxstream_mem_host_deallocate(void* host_mem, libxstream_stream_t* stream);
/** src/libxstream_mem.c:1094-1127 etc.: a NULL stream makes the copy blocking */
LIBXSTREAM_API int libxstream_mem_copy_h2d(const void* host_mem, void* dev_mem, size_t nbytes, libxstream_stream_t* stream);
LIBXSTREAM_API int libxstream_mem_copy_d2h(const void* dev_mem, void* host_mem, size_t nbytes, libxstream_stream_t* stream);
LIBXSTREAM_API int libxstream_mem_copy_d2d(const void* devmem_src, void* devmem_dst, size_t nbytes, libxstream_stream_t* stream);
/** src/libxstream_mem.c:1400-1403: zero nbytes at dev_mem+offset (a fill kernel, not a memset node) */
LIBXSTREAM_API int libxstream_mem_zero(void* dev_mem, size_t offset, size_t nbytes, libxstream_stream_t* stream);

/** libxstream/libxstream_opencl.h:326-330 hint values; src/libxstream_mem.c:401-520 */
typedef enum libxstream_mem_hint_t {
  libxstream_mem_hint_default = 0,
  libxstream_mem_hint_atomics = 1,
  libxstream_mem_hint_compress = 2
} libxstream_mem_hint_t;
/** src/libxstream_mem.c:401: allocation that bypasses any pool; the hint is accepted and ignored */
LIBXSTREAM_API int libxstream_mem_dev_allocate_hint(void** dev_mem, size_t nbytes, libxstream_mem_hint_t hint);
/** src/libxstream_mem.c:488 */
LIBXSTREAM_API int libxstream_mem_dev_deallocate_hint(void* dev_mem);

#endif /*LIBXSTREAM_H*/
