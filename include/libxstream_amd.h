/*
 * libxstream-amd extensions: entry points with no one-to-one counterpart in the reference.
 *
 * They expose what the reference keeps behind environment variables, OpenCL event profiling or
 * LIBXS: explicit kernel configurations (what samples/smm/tune_multiply.py sets through
 * OPENCL_LIBSMM_SMM_* variables, tune_multiply.py:278-298), the per-launch work/duration records of
 * LIBXSTREAM_PROFILE (src/libxstream.c:2629-2781), the tuned-parameter CSV
 * (samples/smm/smm_params.c:51-270) and native HIP handles.
 */
#ifndef LIBXSTREAM_AMD_H
#define LIBXSTREAM_AMD_H

#include "acc_libsmm.h"
#include <stdio.h>

/**
 * Stack-kernel configuration for one (datatype, m, n, k). MI355X-native parameters: the reference's
 * (BS, BM, BN, ...) tuple of samples/smm/smm_acc_opencl.h:60-66 describes OpenCL work-items and is
 * not meaningful for these kernels.
 *   variant: 0 = vector-FMA kernel, operands staged through LDS
 *            1 = matrix-core (MFMA 16x16x4) kernel, operands loaded straight into registers
 *            2 = matrix-core kernel whose wavefronts share each C-run's entries (entry-parallel), their
 *                partial blocks summed in LDS in a fixed order before the one flush of the run
 *            3 = variant 2 on 4x4 blocks (MFMA 4x4x4, four blocks per instruction; FP64 only): C blocks
 *                padded to multiples of 8 instead of 16
 *            4 = 4x4 blocks whose instruction splits k (FP64: four blocks of 4 k, FP32: sixteen blocks of one k,
 *                so 16 of k per instruction for one 4x4 block of C): C blocks up to 16x16 padded to multiples of 4
 *   tm, tn:  variant 0: C microtile (rows x columns) owned by one thread;
 *            variants 1, 2: number of 16x16 output tiles along m and n (padded cover of the C block);
 *            variant 3: m <= 8 tm and n <= 8 tn, the smallest such tm, tn; variant 4: the same with 4
 *   nt:      threads per work-group; variant 0: 64, 128, 256 or 512, the threads left after
 *            covering one C block with microtiles form further groups that share the run's
 *            stack entries; variant 1: 64 per wavefront, the wavefronts split the tile rows;
 *            variants 2-4: 64 per wavefront, the wavefronts split the run's entries
 *   bs:      nominal stack entries per work-group; a work-group owns the C-runs whose first entry
 *            falls in its range and finishes them, so each C block is written by one work-group
 *   kt:      variant 0: k-slice staged in LDS at once (0 = all of k when it fits);
 *            variants 1-4: k per register stage (a multiple of 4 fixed by the kernel; 0 = any)
 *   flush:   0 = automatic, 1 = plain read-modify-write (stack grouped by C), 2 = atomic add.
 *            Configurations reached through libsmm_acc_process (tuned CSV, learned predictor, built-in
 *            default, environment) and the batch always flush with atomic adds, which is correct for
 *            any stack order, unless LIBXSTREAM_SMM_GROUPED=1 declares that every stack holds each C
 *            block in one contiguous run; libsmm_acc_process_config takes flush 1 as that declaration.
 */
typedef struct libsmm_acc_smm_config_t {
  int variant, tm, tn, nt, bs, kt, flush;
  int s;         /* stack size bs was chosen for (CSV column S); shorter stacks launch with bs scaled down
                    (samples/smm/smm_kernel.c:463-473); 0: bs as given */
  int split;     /* C-runs longer than split x bs entries are cut into pieces added atomically (CSV column
                    SPLIT); 0: LIBXSTREAM_SMM_SPLIT (default 2). In stacks of at least 16384 entries runs of
                    up to 64 entries are never cut (bitwise-reproducible C); LIBXSTREAM_SMM_SPLIT_MIN=n
                    sets that floor to n entries for every stack */
  int bs_short, split_short; /* used instead of bs and split for stacks of at most s entries (CSV columns
                    BSS, SPLITS); 0: none. Tuned where they win at s but lose on longer stacks. */
  double gflops; /* measured (autotuner), predicted (learned model) or 0 */
} libsmm_acc_smm_config_t;

/** where a configuration came from */
typedef enum libsmm_acc_config_source_t {
  libsmm_acc_config_default = 0, /* built-in heuristic */
  libsmm_acc_config_tuned = 1,   /* registry: CSV file, libsmm_acc_config_set or autotuner */
  libsmm_acc_config_env = 2,     /* OPENCL_LIBSMM_SMM_* environment overrides */
  libsmm_acc_config_learned = 3  /* learned predictor */
} libsmm_acc_config_source_t;

/** configuration libsmm_acc_process would use for this shape; *source may be NULL */
LIBXSTREAM_API int libsmm_acc_config_get(libsmm_acc_data_t datatype, int m, int n, int k, libsmm_acc_smm_config_t* config,
  int* source);
/** the built-in heuristic's configuration for this shape (ignores the registry and the predictor) */
LIBXSTREAM_API int libsmm_acc_config_heuristic(libsmm_acc_data_t datatype, int m, int n, int k,
  libsmm_acc_smm_config_t* config);
/** register a configuration for this shape (replaces any existing one) */
LIBXSTREAM_API int libsmm_acc_config_set(libsmm_acc_data_t datatype, int m, int n, int k,
  const libsmm_acc_smm_config_t* config);
/** returns EXIT_SUCCESS if config can be launched for this shape (compiled variant, limits) */
LIBXSTREAM_API int libsmm_acc_config_valid(libsmm_acc_data_t datatype, int m, int n, int k,
  const libsmm_acc_smm_config_t* config);
/** libsmm_acc_process with an explicit configuration (the autotuner's entry point) */
LIBXSTREAM_API int libsmm_acc_process_config(const libsmm_acc_smm_config_t* config, const int* dev_param_stack,
  int stack_size, libsmm_acc_data_t datatype, const void* dev_a_data, const void* dev_b_data, void* dev_c_data, int m,
  int n, int k, void* stream);
/** compiled runtime-shape kernels for datatype: returns their number and fills up to cap
 *  (variant, tm, tn, nt, kt) quintuples; kt is the kernel's fixed k per stage (0: any) */
LIBXSTREAM_API int libsmm_acc_kernel_variants(libsmm_acc_data_t datatype, int* quintuples, int cap);
/** every compiled kernel usable for (m, n, k) as (variant, tm, tn, nt, kt) quintuples (kt: the
 *  kernel's fixed k per stage, 0 = any); returns their number, fills up to cap */
LIBXSTREAM_API int libsmm_acc_kernel_candidates(libsmm_acc_data_t datatype, int m, int n, int k, int* quintuples,
  int cap);
/** load tuned parameters from a CSV file (reference or MI355X header); returns #rows loaded or <0. With a
 *  DEVICE column, the rows of the device whose name best matches the active device (share of the row's
 *  name words found in its marketing and gcnArch name) take precedence; rows of other devices fill the
 *  remaining shapes, and among rows of equal standing the higher GFLOPS wins (smm_acc.c:98-235) */
LIBXSTREAM_API int libsmm_acc_params_load(const char* filename);
/** as libsmm_acc_params_load, matched against the given device name instead of the active device's
 *  (NULL: the active device; "": no device matching) */
LIBXSTREAM_API int libsmm_acc_params_load_device(const char* filename, const char* device);
/** write the registry as an MI355X CSV (header line + one row per shape); returns #rows or <0 */
LIBXSTREAM_API int libsmm_acc_params_save(const char* filename, const char* device);

/**
 * Parse one CSV row of the reference's tuned-parameter format (samples/smm/smm_params.c:89-270):
 * [DEVICE;]TYPEID;M;N;K[;S;GFLOPS];BS;BM;BN;BK;WS;WG;LU;NZ;AL;TB;TC;AP;AA;AB;AC[;FLAGS].
 * key[0..3] = type,m,n,k; params[0..15] = bs,bm,bn,bk,ws,wg,lu,nz,al,tb,tc,ap,aa,ab,ac,flags;
 * s_gflops[0..1] = S, GFLOPS. with_device/with_perf select the optional columns.
 * Returns EXIT_SUCCESS when every column parsed.
 */
LIBXSTREAM_API int libsmm_acc_read_ref_params(const char* line, int with_device, int with_perf, char* device,
  size_t device_size, int key[4], int params[16], double s_gflops[2]);

/**
 * Batched stack multiply (no counterpart in the reference): stacks of different (m, n, k) and
 * lengths, each with the semantics of libsmm_acc_process, in one persistent kernel launch; the
 * launch can be captured once into a hipGraph and replayed. m and n must lie in 1..64 (any k);
 * all stacks share the batch's datatype. C-runs are flushed with atomic adds, so stacks may share
 * C pools and need not be sorted.
 */
typedef struct libsmm_acc_batch_stack_t {
  const int* dev_param_stack;
  int stack_size, m, n, k;
  const void* dev_a_data;
  const void* dev_b_data;
  void* dev_c_data;
} libsmm_acc_batch_stack_t;
typedef struct libsmm_acc_batch_t libsmm_acc_batch_t;
/** the batch's schedule: entries per chunk sized to move about chunk_elements operand elements (8..512
 *  entries), blocks_per_cu persistent work-groups per CU; 0 selects the default (32 entries of 23x23x23, 8) */
typedef struct libsmm_acc_batch_params_t {
  int chunk_elements, blocks_per_cu;
  int per_stack; /* 1: one libsmm_acc_process launch per stack (its own configuration) instead of the
                    persistent kernel; libsmm_acc_batch_replay captures all of them in one graph;
                    2: the same launches spread over 4 side streams that fork from the caller's stream and
                    join it again (independent branches of one graph), so that short stacks overlap */
} libsmm_acc_batch_params_t;
/** plan the batch and upload the plan (allocates device memory; not for use under capture) */
LIBXSTREAM_API int libsmm_acc_batch_create(libsmm_acc_data_t datatype, const libsmm_acc_batch_stack_t* stacks,
  int nstacks, libsmm_acc_batch_t** batch);
/** the same with an explicit schedule (NULL: default) */
LIBXSTREAM_API int libsmm_acc_batch_create_ex(libsmm_acc_data_t datatype, const libsmm_acc_batch_stack_t* stacks,
  int nstacks, const libsmm_acc_batch_params_t* params, libsmm_acc_batch_t** batch);
/** autotune the schedule for these stacks: 18 candidates timed on the stacks and pools given, C redirected
 *  to scratch (the caller's C is not touched); blocks until done (allocates; not under capture) */
LIBXSTREAM_API int libsmm_acc_batch_tune(libsmm_acc_data_t datatype, const libsmm_acc_batch_stack_t* stacks,
  int nstacks, libsmm_acc_batch_params_t* best, double* best_ms);
/** enqueue the batch as one kernel launch on stream */
LIBXSTREAM_API int libsmm_acc_batch_process(libsmm_acc_batch_t* batch, void* stream);
/** replay the batch from a hipGraph (captured on the first call) on stream */
LIBXSTREAM_API int libsmm_acc_batch_replay(libsmm_acc_batch_t* batch, void* stream);
LIBXSTREAM_API int libsmm_acc_batch_info(const libsmm_acc_batch_t* batch, int* ntasks, double* flops, double* bytes);
LIBXSTREAM_API int libsmm_acc_batch_destroy(libsmm_acc_batch_t* batch);

/**
 * Learned predictor: an ensemble of FP32 MLPs trained and evaluated on the GPU by the library's own
 * kernels. Where the reference maps an untuned shape to the parameters of the nearest tuned triplet
 * (samples/smm/smm_kernel.c:116-148, :206-213), every network maps (8 shape features: log m, n, k,
 * type, log bytes of the A/B/C blocks, log flops; one-hot encoding of a configuration's variant, tm,
 * tn, nt, kt, bs, flush) through hidden ReLU layers to two outputs: the configuration's throughput
 * relative to the shape's best and log2 of its GFLOP/s. Prediction scores every valid compiled
 * configuration with the ensemble mean and takes the best.
 */
typedef struct libsmm_acc_mlp_t libsmm_acc_mlp_t;
/** nnets members of nhidden layers of width hidden; weights seeded deterministically */
LIBXSTREAM_API int libsmm_acc_mlp_create(libsmm_acc_mlp_t** model, int nnets, int hidden, int nhidden,
  unsigned long long seed);
LIBXSTREAM_API int libsmm_acc_mlp_destroy(libsmm_acc_mlp_t* model);
/** layout of the configuration encoding: seg[p] = first one-hot input of parameter p (after the 8 shape
 *  features; seg[7] = end), values[] = class values; returns the number of inputs */
LIBXSTREAM_API int libsmm_acc_mlp_encoding(int* seg, int* values, int cap);
/** inputs of one (shape, configuration) row; config NULL encodes the shape only */
LIBXSTREAM_API int libsmm_acc_mlp_features(libsmm_acc_data_t datatype, int m, int n, int k,
  const libsmm_acc_smm_config_t* config, float* features);
/** members, layers and widths (width[0] = inputs); returns the parameters per member */
LIBXSTREAM_API int libsmm_acc_mlp_info(const libsmm_acc_mlp_t* model, int* nnets, int* nlayers, int* width, int cap);
/** parameters of member net (copies up to cap when params is not NULL); returns their number */
LIBXSTREAM_API int libsmm_acc_mlp_params(libsmm_acc_mlp_t* model, int net, float* params, int cap);
LIBXSTREAM_API int libsmm_acc_mlp_set_params(libsmm_acc_mlp_t* model, int net, const float* params, int n);
/** one forward and backward pass of member net on the GPU: gradient of the mean weighted squared error
 *  (out_weights per output, NULL = 1) and the loss */
LIBXSTREAM_API int libsmm_acc_mlp_gradient(libsmm_acc_mlp_t* model, int net, const float* x, const float* targets,
  const float* out_weights, int batch, float* grad, double* loss);
/** full-batch Adam on the GPU for epochs steps, every member in turn; mean final loss when loss != NULL */
LIBXSTREAM_API int libsmm_acc_mlp_train(libsmm_acc_mlp_t* model, const float* x, const float* targets,
  const float* out_weights, int batch, int epochs, double lr, double weight_decay, double* loss);
/** inference on the GPU: ensemble mean out[batch][2] */
LIBXSTREAM_API int libsmm_acc_mlp_forward(libsmm_acc_mlp_t* model, const float* x, int batch, float* out);
/** the same inference on the host, with the same per-output operation order (equal results); prediction
 *  (predict_config, predict_topk and the predictor slot of libsmm_acc_process) uses it, so choosing a
 *  configuration allocates no device memory and never waits for the GPU */
LIBXSTREAM_API int libsmm_acc_mlp_forward_host(const libsmm_acc_mlp_t* model, const float* x, int batch, float* out);
LIBXSTREAM_API int libsmm_acc_mlp_save(const libsmm_acc_mlp_t* model, const char* filename);
LIBXSTREAM_API int libsmm_acc_mlp_load(libsmm_acc_mlp_t** model, const char* filename);
/** best valid configuration for the shape by the ensemble (configurations with flush 1 take part only
 *  when LIBXSTREAM_SMM_GROUPED is set, since they launch with the atomic flush otherwise); confidence = its predicted relative
 *  throughput (0..1); config->gflops = predicted GFLOP/s */
LIBXSTREAM_API int libsmm_acc_mlp_predict_config(libsmm_acc_mlp_t* model, libsmm_acc_data_t datatype, int m, int n,
  int k, libsmm_acc_smm_config_t* config, double* confidence);
/** the topk best valid configurations for the shape by the ensemble, best first: scores (may be NULL) =
 *  predicted relative throughput, configs[i].gflops = predicted GFLOP/s; *count = min(topk, valid ones).
 *  The learned-guided autotuner measures only these (tune_multiply -l model -k topk) */
LIBXSTREAM_API int libsmm_acc_mlp_predict_topk(libsmm_acc_mlp_t* model, libsmm_acc_data_t datatype, int m, int n,
  int k, int topk, libsmm_acc_smm_config_t* configs, double* scores, int* count);
/** install (or remove: NULL) the model as predictor for shapes without a tuned configuration; a
 *  prediction below min_confidence leaves the built-in default in place. The caller keeps ownership:
 *  destroy the model only after removing it or after libsmm_acc_finalize */
LIBXSTREAM_API int libsmm_acc_learned_set(libsmm_acc_mlp_t* model, double min_confidence);

/** profiling: per-kernel launch records timed by HIP events on the launch stream */
LIBXSTREAM_API int libxstream_profile_enable(int enable);
LIBXSTREAM_API int libxstream_profile_reset(void);
/** totals for one kernel name ("smm", "trans", ...); waits for outstanding records */
LIBXSTREAM_API int libxstream_profile_get(const char* kernel, long* count, double* total_ms, double* total_flops,
  double* total_bytes);
/** per-name summary of the kept samples: shortest, median and longest duration, the median sample's own
 *  rate (GFLOP/s for kernels, GB/s for transfers, from that sample's work and duration), and the number of
 *  samples discarded as shorter than 10 ticks of the event clock (src/libxstream.c:2640-2660) */
LIBXSTREAM_API int libxstream_profile_stats(const char* name, double* min_ms, double* median_ms, double* max_ms,
  double* median_rate, long* discarded);
/** per-name histograms (LIBXSTREAM_PROFILE=N: N + 1 buckets) as the reference prints them
 *  (src/libxstream.c:853-1015): kernels by duration, transfers (H2D, D2H, D2D, ZERO) by size */
LIBXSTREAM_API int libxstream_profile_print(FILE* stream);
/** copy and fill records (LIBXSTREAM_PROFILE_MEM, src/libxstream_mem.c:805-893) */
LIBXSTREAM_API int libxstream_profile_mem_enable(int enable);

/** last error on the calling thread: returns its code (0 = none) and clears it */
LIBXSTREAM_API int libxstream_error_consume(char* message, size_t size);
/** native handles */
LIBXSTREAM_API int libxstream_stream_native(libxstream_stream_t* stream, void** hip_stream);
/** milliseconds between two recorded events (both must have completed) */
LIBXSTREAM_API int libxstream_event_elapsed(libxstream_event_t* begin, libxstream_event_t* end, double* ms);
/** device properties of the active device: number of CUs, name (gcnArchName) */
LIBXSTREAM_API int libxstream_device_info(int* ncus, char* name, size_t name_size);

#endif /*LIBXSTREAM_AMD_H*/
