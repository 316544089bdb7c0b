/*
 * Launch interface between the C host layer and the gfx950 HIP kernels (C ABI, plain pointers).
 * Every launcher returns 0 on success, a positive hipError_t on a HIP failure, or a negative
 * LXS_E* code when the request cannot be served by a compiled kernel.
 */
#ifndef LXS_KERNELS_H
#define LXS_KERNELS_H

#include <stddef.h>

#if defined(__cplusplus)
extern "C" {
#endif

#define LXS_ENOVARIANT (-2) /* no compiled kernel for (dtype, tm, tn, nt) */
#define LXS_ELIMIT (-3)     /* configuration exceeds a kernel limit (LDS, threads) */

/** data types use the DBCSR codes: 1 = real_4 (float), 3 = real_8 (double) */
typedef struct lxs_smm_args_t {
  void* c;
  const void* a;
  const void* b;
  const int* stack;  /* records of the (a,b,c) element offsets, see pfmt */
  int stack_size;
  int m, n, k;
  int dtype;
  int variant;       /* 0: vector-FMA kernel (LDS staging), 1: MFMA kernel (registers only), 2: entry-parallel
                        MFMA kernel, 3: entry-parallel on 4x4 blocks (FP64), 4: 4x4 blocks splitting k */
  int tm, tn, nt, bs, kt;
  int atomic;        /* 1: flush C-runs with atomic adds; 0: plain read-modify-write */
  int pfmt;          /* 0: 3-int records, 1-based offsets; else 6-int DBM task records (m, n, k, a, b, c),
                        0-based offsets (samples/smm/kernels/multiply.cl:96-99) */
  int tsplit;        /* C-runs longer than this are cut at segment boundaries (<= 0: never) */
} lxs_smm_args_t;

/** number of LDS bytes the stack kernel needs for this configuration (0 if invalid) */
/* 1 when the configuration can be launched (LDS fits; wavefront-per-run kernels need none) */
int lxs_smm_launchable(const lxs_smm_args_t* args);
size_t lxs_smm_lds_bytes(const lxs_smm_args_t* args);
/** 1 if a kernel is compiled for (dtype, variant, tm, tn, nt, kt) and shape, 2 if specialised for
 *  (m, n, k), else 0 */
int lxs_smm_available(int dtype, int variant, int m, int n, int k, int tm, int tn, int nt, int kt);
/** enumerate compiled runtime-shape kernels for dtype: returns count, fills up to cap quintuples
 *  (variant, tm, tn, nt, kt) where kt is the compile-time k per stage (0: any) */
int lxs_smm_variants(int dtype, int* quintuples, int cap);
/** every compiled kernel usable for (m, n, k) as (variant, tm, tn, nt, kt); returns the count */
int lxs_smm_candidates(int dtype, int m, int n, int k, int* quintuples, int cap);
/** first compiled kernel for exactly this shape (specialised), else -1; fills a quintuple */
int lxs_smm_specialised(int dtype, int variant, int m, int n, int k, int* quintuple);
/** largest LDS allocation one work-group may use */
size_t lxs_lds_limit(void);
int lxs_smm_launch(const lxs_smm_args_t* args, void* hip_stream);

/** one stack of a batch (device pointers; m, n <= 64) */
typedef struct lxs_batch_stack_t {
  const void* a;
  const void* b;
  void* c;
  const int* stack;
  int stack_size, m, n, k;
  int bs;     /* entries per chunk */
  int ti, tj; /* 16x16 output tiles covering the C block */
  int tsplit; /* C-runs longer than this are cut at chunk boundaries (<= 0: never) */
} lxs_batch_stack_t;

/** up to 4 chunks (one per wavefront), or one chunk split over the 4 wavefronts (team != 0) */
typedef struct lxs_batch_task_t {
  int stack[4]; /* stack index, -1 = none */
  int chunk[4];
  int team;
  int pad[3];
} lxs_batch_task_t;

int lxs_batch_launch(int dtype, const lxs_batch_stack_t* desc, const lxs_batch_task_t* tasks, int ntasks, int nblocks,
  void* hip_stream);

int lxs_trans_launch(const int* trs_stack, int offset, int stack_size, void* data, int dtype, int m, int n,
  void* hip_stream);
int lxs_zero_launch(void* ptr, size_t nbytes, void* hip_stream);

/* Ozaki scheme I (ozaki.hip). Operand element (r, k) at p[r * rs + k * ks] (rs or ks is 1);
   planes: s x Rp x Kp int8 (Rp multiple of 256, Kp multiple of 128); e: Rp ints; occ: 8 ints */
int lxs_oz_preprocess(int dtype, const void* p, long rs, long ks, int R, int K, int s, int Rp, int Kp, int* e,
  int* occ, signed char* planes, void* hip_stream);
/* int8 GEMM over work items {level, u0, u1, pad} (host array), element (i, j) of item z at
   out + z Mp Np + j Mp + i. crt = 0 (scheme 1): s digit planes, item level L over the pairs (p, L - p),
   int32 partials, at most 2080 units (pair, k-block of 64) per item. crt = 1 (scheme 2): s residue
   planes, item level = prime, sums mod m_prime as bytes, at most 2047 units (k-blocks) per item. */
int lxs_oz_level_gemm(int crt, const signed char* pa, const signed char* pb, const int* occa, const int* occb,
  const int* items, int nitems, int s, int Mp, int Np, int Kp, void* out, void* hip_stream);
/* FP64 recombination of the partials: zb[L] = first item of level L (cutoff + 2 entries) */
int lxs_oz_recombine(int dtype, const int* c32, const int* zb, int cutoff, int Mp, int Np, const int* ea,
  const int* eb, int M, int N, double alpha, double beta, void* c, int ldc, void* hip_stream);
int lxs_oz_scale(int dtype, void* c, int ldc, int M, int N, double beta, void* hip_stream);
/* bytes of the per-device static scratch of ozaki_gemm_noalloc (ozaki.hip) */
#define LXS_OZ_STATIC_BYTES ((size_t)64 << 20)
/* device address of that scratch on the current device (looked up once per device): 0, a hipError_t or
   LXS_ELIMIT */
int lxs_oz_static_scratch(void** ptr);
/* Ozaki scheme 2 (CRT): moduli in use order, pairwise coprime and at most 256, so that symmetric residues
   fit int8; the first np of them serve a call. Residue planes: np x Rp x Kp int8 of the integers
   trunc(|v| 2^(mb - trim - E)) (E: the row's exponent, mb = 53 or 24) with the sign of v, times the
   reconstruction weight w_p = ((M / m_p) mod m_p)^-1 mod m_p when weight != 0 (one operand carries it, so
   the int8 GEMM's residues are the CRT coefficients). */
#define LXS_OZ_NPRIMES_MAX 16
#define LXS_OZ_MODULI 256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 211, 199, 197, 193, 191
int lxs_oz_crt_preprocess(int dtype, const void* p, long rs, long ks, int R, int K, int np, int weight, int trim, int Rp,
  int Kp, int* e, signed char* planes, void* hip_stream);
/* C = alpha 2^(ea + eb + sh0) x + beta C, x the integer (|x| < M / 2, M the product of the first np moduli) whose
   CRT coefficients c_p = x w_p mod m_p are the sums of the residue bytes of the prime's items zb[p] .. zb[p + 1] - 1,
   mod m_p (the A operand's residues carry w_p) */
int lxs_oz_crt_recombine(int dtype, const unsigned char* r8, const int* zb, int np, int Mp, int Np, const int* ea,
  const int* eb, int M, int N, int sh0, double alpha, double beta, void* c, int ldc, void* hip_stream);
int lxs_dgemm(const double* a, long ars, long aks, const double* b, long bks, long bcs, int M, int N, int K,
  double alpha, double beta, double* c, int ldc, void* hip_stream);
/* complex 3M helpers: split interleaved (rows x cols, ld) into re, im, re + im (ld rows), conj negates im;
   combine C = alpha (t1 - t2 + i (t3 - t1 - t2)) + beta C */
int lxs_oz_csplit(int dtype, const void* z, int ld, int rows, int cols, int conj, void* re, void* im, void* sum,
  void* hip_stream);
int lxs_oz_ccombine(int dtype, const void* t1, const void* t2, const void* t3, int M, int N, const double* alpha,
  const double* beta, void* c, int ldc, void* hip_stream);

/* Ozaki scheme I over a parameter stack (ozaki_stack.hip): planes_a na * s * 4 KB, planes_b nb * s * 4 KB,
   ea na * 64 ints, eb nb * 64 ints, info 1 int of device scratch */
int lxs_ozs_process(int dtype, const int* stack, int S, int m, int n, int k, const void* a, int na, const void* b,
  int nb, void* c, int s, int cutoff, signed char* planes_a, int* ea, signed char* planes_b, int* eb, int* info,
  void* hip_stream);

/* learned predictor MLP (mlp.hip): activations [batch][features], weights [out][in] then bias [out] */
int lxs_mlp_forward(const float* x, const float* w, int batch, int nin, int nout, int relu, float* y, void* hip_stream);
int lxs_mlp_loss(const float* z, const float* tgt, const float* hw, int batch, int nh, float* dz, float* loss,
  void* hip_stream);
/* scratch (floats) of lxs_mlp_backward; layers up to 64 outputs x 64 inputs */
long lxs_mlp_scratch(int batch, int nin, int nout);
int lxs_mlp_backward(const float* x, const float* y, const float* w, float* dz, int batch, int nin, int nout, int relu,
  float* dw, float* dx, float* scratch, void* hip_stream);
int lxs_mlp_adam(float* p, const float* g, float* m, float* v, int n, float lr, float beta1, float beta2, float eps,
  int step, float wd, void* hip_stream);

#if defined(__cplusplus)
}
#endif

#endif /*LXS_KERNELS_H*/
