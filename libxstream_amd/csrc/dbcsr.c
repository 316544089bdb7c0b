/*
 * DBCSR accelerator interface: c_dbcsr_acc_* over the libxstream runtime.
 * One wrapper per function of src/libxstream_dbcsr.c:37-304, same argument forwarding. With
 * -DLIBXSTREAM_DBCSR_TIMERS each call is bracketed by DBCSR's c_dbcsr_timeset/timestop
 * (libxstream_dbcsr.c:15-34) when LIBXSTREAM_PROFILE is set.
 */
#include "lxs_internal.h"

#if defined(LIBXSTREAM_DBCSR_TIMERS)
#  define LXS_TIMER_BEGIN \
    int lxs_timer_handle_ = 0; \
    static const char* const lxs_timer_name_ = __func__; \
    static const int lxs_timer_len_ = (int)sizeof(__func__) - 1; \
    if (0 != lxs_cfg.profile) c_dbcsr_timeset(&lxs_timer_name_, &lxs_timer_len_, &lxs_timer_handle_)
#  define LXS_TIMER_END \
    if (0 != lxs_cfg.profile) c_dbcsr_timestop(&lxs_timer_handle_)
#else
#  define LXS_TIMER_BEGIN
#  define LXS_TIMER_END
#endif

#define LXS_WRAP(EXPR) \
  int result; \
  LXS_TIMER_BEGIN; \
  result = (EXPR); \
  LXS_TIMER_END; \
  return result

LIBXSTREAM_API int c_dbcsr_acc_init(void) { LXS_WRAP(libxstream_init()); }
LIBXSTREAM_API int c_dbcsr_acc_finalize(void) { LXS_WRAP(libxstream_finalize()); }
LIBXSTREAM_API void c_dbcsr_acc_clear_errors(void) { (void)libxstream_error_consume(NULL, 0); }
LIBXSTREAM_API int c_dbcsr_acc_get_ndevices(int* ndevices) { LXS_WRAP(libxstream_device_count(ndevices)); }
LIBXSTREAM_API int c_dbcsr_acc_set_active_device(int device_id) { LXS_WRAP(libxstream_device_set_active(device_id)); }
LIBXSTREAM_API int c_dbcsr_acc_device_synchronize(void) { LXS_WRAP(libxstream_device_sync()); }

LIBXSTREAM_API int c_dbcsr_acc_stream_priority_range(int* least, int* greatest) {
  LXS_WRAP(libxstream_stream_priority_range(least, greatest));
}

/* libxstream_dbcsr.c:105-124: priority maps to LOW/HIGH flags only when a range exists; there is none */
LIBXSTREAM_API int c_dbcsr_acc_stream_create(void** stream_p, const char* name, int priority) {
  (void)priority;
  LXS_WRAP(libxstream_stream_create((libxstream_stream_t**)stream_p, name, LIBXSTREAM_STREAM_DEFAULT));
}

LIBXSTREAM_API int c_dbcsr_acc_stream_destroy(void* stream) {
  LXS_WRAP(libxstream_stream_destroy((libxstream_stream_t*)stream));
}
LIBXSTREAM_API int c_dbcsr_acc_stream_sync(void* stream) { LXS_WRAP(libxstream_stream_sync((libxstream_stream_t*)stream)); }
LIBXSTREAM_API int c_dbcsr_acc_stream_wait_event(void* stream, void* event) {
  LXS_WRAP(libxstream_stream_wait_event((libxstream_stream_t*)stream, (libxstream_event_t*)event));
}

LIBXSTREAM_API int c_dbcsr_acc_event_create(void** event_p) {
  LXS_WRAP(libxstream_event_create((libxstream_event_t**)event_p));
}
LIBXSTREAM_API int c_dbcsr_acc_event_destroy(void* event) { LXS_WRAP(libxstream_event_destroy((libxstream_event_t*)event)); }
LIBXSTREAM_API int c_dbcsr_acc_event_record(void* event, void* stream) {
  LXS_WRAP(libxstream_event_record((libxstream_event_t*)event, (libxstream_stream_t*)stream));
}
LIBXSTREAM_API int c_dbcsr_acc_event_query(void* event, c_dbcsr_acc_bool_t* has_occurred) {
  LXS_WRAP(libxstream_event_query((libxstream_event_t*)event, has_occurred));
}
LIBXSTREAM_API int c_dbcsr_acc_event_synchronize(void* event) {
  LXS_WRAP(libxstream_event_sync((libxstream_event_t*)event));
}

LIBXSTREAM_API int c_dbcsr_acc_dev_mem_allocate(void** dev_mem, size_t nbytes) {
  LXS_WRAP(libxstream_mem_allocate(dev_mem, nbytes));
}
LIBXSTREAM_API int c_dbcsr_acc_dev_mem_deallocate(void* dev_mem) { LXS_WRAP(libxstream_mem_deallocate(dev_mem)); }
LIBXSTREAM_API int c_dbcsr_acc_dev_mem_set_ptr(void** dev_mem, void* other, size_t lb) {
  LXS_WRAP(libxstream_mem_offset(dev_mem, other, lb));
}
LIBXSTREAM_API int c_dbcsr_acc_host_mem_allocate(void** host_mem, size_t nbytes, void* stream) {
  LXS_WRAP(libxstream_mem_host_allocate(host_mem, nbytes, (libxstream_stream_t*)stream));
}
LIBXSTREAM_API int c_dbcsr_acc_host_mem_deallocate(void* host_mem, void* stream) {
  LXS_WRAP(libxstream_mem_host_deallocate(host_mem, (libxstream_stream_t*)stream));
}
LIBXSTREAM_API int c_dbcsr_acc_memcpy_h2d(const void* host_mem, void* dev_mem, size_t nbytes, void* stream) {
  LXS_WRAP(libxstream_mem_copy_h2d(host_mem, dev_mem, nbytes, (libxstream_stream_t*)stream));
}
LIBXSTREAM_API int c_dbcsr_acc_memcpy_d2h(const void* dev_mem, void* host_mem, size_t nbytes, void* stream) {
  LXS_WRAP(libxstream_mem_copy_d2h(dev_mem, host_mem, nbytes, (libxstream_stream_t*)stream));
}
LIBXSTREAM_API int c_dbcsr_acc_memcpy_d2d(const void* devmem_src, void* devmem_dst, size_t nbytes, void* stream) {
  LXS_WRAP(libxstream_mem_copy_d2d(devmem_src, devmem_dst, nbytes, (libxstream_stream_t*)stream));
}
LIBXSTREAM_API int c_dbcsr_acc_memset_zero(void* dev_mem, size_t offset, size_t nbytes, void* stream) {
  LXS_WRAP(libxstream_mem_zero(dev_mem, offset, nbytes, (libxstream_stream_t*)stream));
}
LIBXSTREAM_API int c_dbcsr_acc_dev_mem_info(size_t* mem_free, size_t* mem_total) {
  LXS_WRAP(libxstream_mem_info(mem_free, mem_total));
}
