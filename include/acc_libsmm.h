/*
 * libxstream-amd: ACC LIBSMM interface (libsmm_acc_*), the hot-path boundary DBCSR calls.
 *
 * Replaces samples/smm/acc_libsmm.h:31-57 (hfp/libxstream). Signatures, datatype codes and the
 * return convention are those of the reference:
 *   - libsmm_acc_process returns -1 when the stack is not handled and the caller must fall back
 *     to a host multiply (samples/smm/smm_kernel.c:524-526); 0 on success; any other non-zero
 *     value is a runtime error.
 *   - stack_size == 0 is success (nothing to do).
 * The kernels behind these entry points are hand-written HIP for gfx950 (MI355X).
 */
#ifndef DBCSR_ACC_LIBSMM_H
#define DBCSR_ACC_LIBSMM_H

#include "libxstream_dbcsr.h"

#define DBCSR_TYPE(T) DBCSR_CONCATENATE(DBCSR_TYPE_, T)
#define DBCSR_TYPE_double dbcsr_type_real_8
#define DBCSR_TYPE_float dbcsr_type_real_4

/** samples/smm/acc_libsmm.h:31-36 */
typedef enum libsmm_acc_data_t {
  dbcsr_type_real_4 = 1,
  dbcsr_type_real_8 = 3,
  dbcsr_type_complex_4 = 5,
  dbcsr_type_complex_8 = 7
} libsmm_acc_data_t;

/** samples/smm/acc_libsmm.h:38; reference-counted, may initialize libxstream (smm_acc.c:43-62) */
LIBXSTREAM_API int libsmm_acc_init(void);
/** samples/smm/acc_libsmm.h:39 (smm_acc.c:427-481) */
LIBXSTREAM_API int libsmm_acc_finalize(void);
/** samples/smm/acc_libsmm.h:40 (smm_acc.c:484-491): concurrent process/transpose calls are safe */
LIBXSTREAM_API c_dbcsr_acc_bool_t libsmm_acc_is_thread_safe(void);

/**
 * samples/smm/acc_libsmm.h:42-43 (smm_trans.c:48-228, kernels/transpose.cl:15-84).
 * In place, for i in [0, stack_size): the m-by-n column-major block at
 * dev_data + dev_trs_stack[offset + i] (0-based element offset) becomes row-major, i.e. the
 * n-by-m column-major transpose. Unsupported datatype, m*n == 1 or m*n > max_kernel_dim^2 is a
 * successful no-op, as in the reference.
 */
LIBXSTREAM_API int libsmm_acc_transpose(const int* dev_trs_stack, int offset, int stack_size, void* dev_data,
  libsmm_acc_data_t datatype, int m, int n, int max_kernel_dim, void* stream);

/**
 * samples/smm/acc_libsmm.h:45-47 (smm_acc.c:575-593, smm_kernel.c:82-528, kernels/multiply.cl).
 * For i in [0, stack_size), with (a,b,c) = dev_param_stack[3i .. 3i+2] (1-based element offsets):
 *   C[c-1 ...] (m-by-n, column-major) += A[a-1 ...] (m-by-k, column-major) * B[b-1 ...]
 *   where the B block is stored row-major k-by-n (the layout libsmm_acc_transpose produces).
 * Asynchronous on stack_stream; host_param_stack and c_stream are not used (smm_kernel.c:86-87).
 * Returns -1 (host fallback) unless m*n > 0, k > 0, def_mnk != 0, m*n <= max_kernel_dim^2 and
 * datatype is real_8 or real_4 (smm_acc.c:527-573).
 */
LIBXSTREAM_API int libsmm_acc_process(const int* host_param_stack, const int* dev_param_stack, int stack_size,
  libsmm_acc_data_t datatype, const void* dev_a_data, const void* dev_b_data, void* dev_c_data, int m_max, int n_max,
  int k_max, int max_kernel_dim, c_dbcsr_acc_bool_t def_mnk, void* stack_stream, void* c_stream);

/** samples/smm/acc_libsmm.h:49 (smm_acc.c:596-604): not implemented by the reference either, returns -1 */
LIBXSTREAM_API int c_calculate_norms(const double* mat, int nblks, const int* offsets, const int* nelems, float* norms,
  void* stream_ptr);

/** samples/smm/smm_acc_opencl.h:111-112 (smm_acc.c:527-573): the suitability predicate behind the -1 result */
LIBXSTREAM_API c_dbcsr_acc_bool_t libsmm_acc_process_suitable(c_dbcsr_acc_bool_t def_mnk, libsmm_acc_data_t datatype,
  int stack_size, int m_max, int n_max, int k_max, int max_kernel_dim);

/**
 * DBM reuse of the stack kernels (samples/smm/smm_acc_opencl.h:29-31,114-124; smm_acc.c:579-585;
 * smm_kernel.c:73-76). Bits per extent when libsmm_acc_process packs a shape into param_format.
 */
#define LIBSMM_ACC_PFORMAT 8

/** samples/smm/smm_acc_opencl.h:115-116: CP2K's DBM launcher (shard_c_data += pack_a * pack_b per task) */
typedef int (*opencl_libsmm_acc_dbm_launch_fn_t)(void* stream, double alpha, int ntasks, int param_format,
  const int* params_host, const int* params, const double* pack_a_data, const double* pack_b_data,
  double* shard_c_data);

/**
 * samples/smm/smm_acc_opencl.h:117-118 (smm_kernel.c:74-76). Once a function is registered (NULL
 * unregisters), libsmm_acc_process hands every real_8 stack with def_mnk != 0 and m, n, k < 128 to it,
 * with alpha 1, param_format = m | n << 8 | k << 16, and returns its result (smm_acc.c:579-585).
 */
LIBXSTREAM_API void opencl_libsmm_acc_set_dbm_launch_fn(opencl_libsmm_acc_dbm_launch_fn_t launch_fn);

/**
 * samples/smm/smm_acc_opencl.h:120-123 (smm_kernel.c:82-528): libsmm_acc_process over records of
 * either format, so DBM tasks run on the same kernels.
 *   - param_format == 0: records of 3 ints (a, b, c), 1-based element offsets (as libsmm_acc_process).
 *   - param_format != 0: DBM task records of 6 ints (m, n, k, a, b, c), 0-based element offsets; m, n, k
 *     of a record are not read: every task has the shape m_max x n_max x k_max (multiply.cl:96-99).
 * The result convention is libsmm_acc_process's (-1: host fallback). event: NULL, or an event from
 * c_dbcsr_acc_event_create, recorded on stream after the launch (in place of the reference's cl_event).
 */
LIBXSTREAM_API int opencl_libsmm_acc_process(const int* host_param_stack, const int* dev_param_stack, int stack_size,
  libsmm_acc_data_t datatype, const void* dev_a_data, const void* dev_b_data, void* dev_c_data, int m_max, int n_max,
  int k_max, int max_kernel_dim, c_dbcsr_acc_bool_t def_mnk, void* stream, void* c_stream, int param_format,
  void* event);

/** DBCSR timer names (samples/smm/acc_libsmm.h:51-57) */
static const char libsmm_acc_transpose_routine_name_str[] = "jit_kernel_transpose";
static const char libsmm_acc_process_routine_name_str[] = "jit_kernel_multiply";

#endif /*DBCSR_ACC_LIBSMM_H*/
