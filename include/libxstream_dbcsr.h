/*
 * libxstream-amd: DBCSR accelerator interface (c_dbcsr_acc_*), the drop-in boundary DBCSR binds.
 *
 * Replaces libxstream/libxstream_dbcsr.h:24-67 and the wrappers in src/libxstream_dbcsr.c:37-304
 * (hfp/libxstream). Every function keeps the reference signature; handles are opaque void* on the
 * DBCSR side and libxstream_stream_t* / libxstream_event_t* inside. Return codes follow
 * libxstream.h (0 = success).
 */
#ifndef LIBXSTREAM_DBCSR_H
#define LIBXSTREAM_DBCSR_H

#include "libxstream.h"

#define DBCSR_STRINGIFY_AUX(SYMBOL) #SYMBOL
#define DBCSR_STRINGIFY(SYMBOL) DBCSR_STRINGIFY_AUX(SYMBOL)
#define DBCSR_CONCATENATE2(A, B) A##B
#define DBCSR_CONCATENATE(A, B) DBCSR_CONCATENATE2(A, B)
#define DBCSR_MARK_USED(x) (void)(x)

/** types (libxstream/libxstream_dbcsr.h:24) */
typedef int c_dbcsr_acc_bool_t;

/** initialization and finalization (libxstream_dbcsr.h:27-28; libxstream_dbcsr.c:37-54) */
LIBXSTREAM_API int c_dbcsr_acc_init(void);
LIBXSTREAM_API int c_dbcsr_acc_finalize(void);

/** error handling (libxstream_dbcsr.h:31; libxstream_dbcsr.c:57-62): consumes the last error */
LIBXSTREAM_API void c_dbcsr_acc_clear_errors(void);

/** devices (libxstream_dbcsr.h:34-36; libxstream_dbcsr.c:65-92) */
LIBXSTREAM_API int c_dbcsr_acc_get_ndevices(int* ndevices);
LIBXSTREAM_API int c_dbcsr_acc_set_active_device(int device_id);
LIBXSTREAM_API int c_dbcsr_acc_device_synchronize(void);

/** streams (libxstream_dbcsr.h:39-45; libxstream_dbcsr.c:95-154) */
LIBXSTREAM_API int c_dbcsr_acc_stream_priority_range(int* least, int* greatest);
LIBXSTREAM_API int c_dbcsr_acc_stream_create(void** stream_p, const char* name,
  /** lower number is higher priority */
  int priority);
LIBXSTREAM_API int c_dbcsr_acc_stream_destroy(void* stream);
LIBXSTREAM_API int c_dbcsr_acc_stream_sync(void* stream);
LIBXSTREAM_API int c_dbcsr_acc_stream_wait_event(void* stream, void* event);

/** events (libxstream_dbcsr.h:48-52; libxstream_dbcsr.c:157-204) */
LIBXSTREAM_API int c_dbcsr_acc_event_create(void** event_p);
LIBXSTREAM_API int c_dbcsr_acc_event_destroy(void* event);
LIBXSTREAM_API int c_dbcsr_acc_event_record(void* event, void* stream);
LIBXSTREAM_API int c_dbcsr_acc_event_query(void* event, c_dbcsr_acc_bool_t* has_occurred);
LIBXSTREAM_API int c_dbcsr_acc_event_synchronize(void* event);

/** memory (libxstream_dbcsr.h:55-64; libxstream_dbcsr.c:207-304) */
LIBXSTREAM_API int c_dbcsr_acc_dev_mem_allocate(void** dev_mem, size_t nbytes);
LIBXSTREAM_API int c_dbcsr_acc_dev_mem_deallocate(void* dev_mem);
LIBXSTREAM_API int c_dbcsr_acc_dev_mem_set_ptr(void** dev_mem, void* other, size_t lb);
LIBXSTREAM_API int c_dbcsr_acc_host_mem_allocate(void** host_mem, size_t nbytes, void* stream);
LIBXSTREAM_API int c_dbcsr_acc_host_mem_deallocate(void* host_mem, void* stream);
LIBXSTREAM_API int c_dbcsr_acc_memcpy_h2d(const void* host_mem, void* dev_mem, size_t nbytes, void* stream);
LIBXSTREAM_API int c_dbcsr_acc_memcpy_d2h(const void* dev_mem, void* host_mem, size_t nbytes, void* stream);
LIBXSTREAM_API int c_dbcsr_acc_memcpy_d2d(const void* devmem_src, void* devmem_dst, size_t nbytes, void* stream);
LIBXSTREAM_API int c_dbcsr_acc_memset_zero(void* dev_mem, size_t offset, size_t nbytes, void* stream);
LIBXSTREAM_API int c_dbcsr_acc_dev_mem_info(size_t* mem_free, size_t* mem_total);

/** imported from DBCSR, not exported (libxstream_dbcsr.h:66-67); used only when built with
 *  -DLIBXSTREAM_DBCSR_TIMERS, which brackets every c_dbcsr_acc_* call like libxstream_dbcsr.c:15-34 */
void c_dbcsr_timeset(const char** routineN, const int* routineN_len, int* handle);
void c_dbcsr_timestop(const int* handle);

#endif /*LIBXSTREAM_DBCSR_H*/
