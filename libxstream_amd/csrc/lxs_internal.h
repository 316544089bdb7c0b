/*
 * Internal declarations shared by the C host layer (runtime, DBCSR shim, LIBSMM host code).
 */
#ifndef LXS_INTERNAL_H
#define LXS_INTERNAL_H

#include <hip/hip_runtime_api.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include "../../include/libxstream_amd.h"
#include "lxs_kernels.h"

struct libxstream_stream_t {
  hipStream_t hip;
  int device;
  int flags; /* libxstream_stream_flags_t as requested */
  char name[32];
};

struct libxstream_event_t {
  hipEvent_t hip;
};

typedef struct lxs_config_t {
  int ninit;     /* reference count of libxstream_init */
  int ndevices;  /* 0 when no HIP device is visible */
  int device;    /* active device, -1 if none */
  int verbosity;
  int profile;   /* LIBXSTREAM_PROFILE: launch records (|value| + 1 histogram buckets; < 0 also prints each) */
  int profile_mem; /* LIBXSTREAM_PROFILE_MEM: copy and fill records (H2D, D2H, D2D, ZERO) */
  int split;     /* LIBXSTREAM_SMM_SPLIT: C-runs are cut past max(split*bs, split_min) entries (0: never) */
  int split_min; /* LIBXSTREAM_SMM_SPLIT_MIN: runs of up to this many entries are never cut (0: 64 in stacks of
                    at least LXS_SMM_SPLIT_FILL entries, none in shorter ones) */
  int grouped;   /* LIBXSTREAM_SMM_GROUPED: the caller guarantees that every stack holds each C block in one
                    contiguous run, which lets configurations with flush 1 write C without atomics (default 0) */
  int ncus;      /* compute units of the active device */
  char arch[256]; /* gcnArchName of the active device */
  char devname[512]; /* marketing name and gcnArchName of the active device ("" before one is active) */
  pthread_mutex_t lock;
} lxs_config_t;

extern lxs_config_t lxs_cfg;

/** record code/message as the calling thread's last error; prints at verbosity >= 1; returns code */
int lxs_error(int code, const char* what);
/** map a hipError_t to the library's return convention (0 or the positive HIP code) */
int lxs_hip(hipError_t code, const char* what);
/** bind the calling thread to the active device (HIP keeps the current device per thread) */
int lxs_ensure_device(void);
/** HIP stream behind a library stream; NULL selects the calling thread's default stream */
hipStream_t lxs_stream_hip(const libxstream_stream_t* stream);

/** profiling around one launch: token is NULL when profiling is off */
void* lxs_profile_begin(hipStream_t stream);
void* lxs_profile_begin_mem(hipStream_t stream);
void lxs_profile_end(void* token, hipStream_t stream, const char* kernel, double flops, double bytes);
void lxs_profile_finalize(void);

/** learned predictor slot (mlp.c): EXIT_SUCCESS and a configuration when a confident model is installed */
int lxs_learned_predict(int datatype, int m, int n, int k, libsmm_acc_smm_config_t* config);
void lxs_learned_finalize(void);
/** load a model file and install it as the predictor, owned by the library until finalize */
void lxs_learned_load_default(const char* path);

/** LIBSMM registry lifetime (smm.c) */
int lxs_smm_init(void);
void lxs_smm_finalize(void);
/* C-runs of a launch with bs entries per work-group are cut past this many entries (0: never);
   split: the configuration's factor (0: the library default); stacks of at least LXS_SMM_SPLIT_FILL
   entries keep runs of up to 64 entries whole (unless LIBXSTREAM_SMM_SPLIT_MIN says otherwise) */
#define LXS_SMM_SPLIT_FILL 16384
int lxs_smm_tsplit(int bs, int split, int stack_size);
/* one stack with its resolved configuration on a HIP stream, no profile records (hipError_t or LXS_E*) */
int lxs_smm_enqueue(int datatype, const int* dev_param_stack, int stack_size, const void* a, const void* b, void* c,
  int m, int n, int k, void* hip_stream);

#endif /*LXS_INTERNAL_H*/
