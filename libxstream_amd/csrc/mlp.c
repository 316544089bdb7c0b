/*
 * Learned stack-kernel predictor: an ensemble of FP32 MLPs trained and evaluated with the HIP kernels
 * of mlp.hip.
 *
 * The reference picks parameters for an untuned (m, n, k) from the nearest tuned triplet
 * (samples/smm/smm_kernel.c:116-148, :206-213) or falls back to defaults. Here every network maps
 * (8 shape features, the one-hot encoding of a configuration's parameters: variant, tm, tn, nt, kt,
 * bs, flush) through 3 ReLU layers to two outputs: the configuration's throughput relative to the
 * best one for the shape, and log2 of its GFLOP/s. Every compiled configuration that is valid for the
 * shape is scored with the ensemble mean of the first output; the best one is taken when its
 * predicted relative throughput reaches the predictor's confidence threshold, otherwise the default
 * heuristic stays (and the exhaustive autotuner, tune_multiply, remains the fallback).
 */
#include "lxs_internal.h"
#include <math.h>
#include <stdint.h>
#include <string.h>

#define MLP_NSHAPE 8  /* shape features */
#define MLP_NPARAM 7  /* encoded configuration parameters */
#define MLP_NOUT 2    /* relative throughput, log2 GFLOP/s */
#define MLP_MAXL 8
#define MLP_MAXNETS 16

/* class values of the encoded parameters (one-hot) */
static const int mlp_variant[] = {0, 1, 2, 3, 4};
static const int mlp_tm[] = {1, 2, 3, 4, 5};
static const int mlp_tn[] = {1, 2, 3, 4, 5};
static const int mlp_nt[] = {64, 128, 192, 256, 320, 512};
static const int mlp_kt[] = {0, 8, 12, 16, 24, 32};
static const int mlp_bs[] = {16, 32, 64, 128};
static const int mlp_flush[] = {1, 2};
static const int* const mlp_vals[MLP_NPARAM] = {mlp_variant, mlp_tm, mlp_tn, mlp_nt, mlp_kt, mlp_bs, mlp_flush};
static const int mlp_nvals[MLP_NPARAM] = {5, 5, 5, 6, 6, 4, 2};

struct libsmm_acc_mlp_t {
  int nnets;                 /* ensemble members, identical shape */
  int nlayers;               /* weight layers per network */
  int width[MLP_MAXL + 1];   /* width[0] = inputs, width[nlayers] = outputs */
  int nparams;               /* per network */
  float* params;             /* host copy, nnets x nparams (authoritative between calls) */
  float *d_params, *d_grad, *d_m, *d_v; /* device state, nnets x nparams each */
  int step[MLP_MAXNETS];
};

static libsmm_acc_mlp_t* lxs_learned = NULL;
static libsmm_acc_mlp_t* lxs_learned_owned = NULL; /* the shipped model, loaded at init */
static double lxs_learned_conf = 0.0;
#define MLP_DEFAULT_CONF 0.0 /* confidence threshold of the shipped model (LIBXSTREAM_LEARNED_CONF) */
static pthread_mutex_t lxs_learned_lock = PTHREAD_MUTEX_INITIALIZER;

static int mlp_nenc(void) {
  int h, n = 0;
  for (h = 0; h < MLP_NPARAM; ++h) n += mlp_nvals[h];
  return n;
}

LIBXSTREAM_API int libsmm_acc_mlp_encoding(int* seg, int* values, int cap) {
  int h, n = 0;
  for (h = 0; h < MLP_NPARAM; ++h) {
    int v;
    if (NULL != seg && h < cap) seg[h] = n;
    for (v = 0; v < mlp_nvals[h]; ++v) {
      if (NULL != values && n + v < cap) values[n + v] = mlp_vals[h][v];
    }
    n += mlp_nvals[h];
  }
  if (NULL != seg && MLP_NPARAM < cap) seg[MLP_NPARAM] = n;
  return MLP_NSHAPE + n;
}

static int mlp_class(int param, int value) {
  int v;
  for (v = 0; v < mlp_nvals[param]; ++v) {
    if (mlp_vals[param][v] == value) return v;
  }
  return -1;
}

LIBXSTREAM_API int libsmm_acc_mlp_features(libsmm_acc_data_t datatype, int m, int n, int k,
  const libsmm_acc_smm_config_t* config, float* feat) {
  const double ts = (dbcsr_type_real_8 == datatype ? 8.0 : (dbcsr_type_real_4 == datatype ? 4.0 : 0.0));
  int h, off = MLP_NSHAPE;
  if (NULL == feat || 0 >= m || 0 >= n || 0 >= k || 0 == ts) return EXIT_FAILURE;
  feat[0] = (float)((log2((double)m) - 4.0) / 1.5);
  feat[1] = (float)((log2((double)n) - 4.0) / 1.5);
  feat[2] = (float)((log2((double)k) - 4.0) / 1.5);
  feat[3] = (8.0 == ts ? 1.f : -1.f);
  feat[4] = (float)((log2(ts * m * k) - 12.0) / 2.5);
  feat[5] = (float)((log2(ts * k * n) - 12.0) / 2.5);
  feat[6] = (float)((log2(ts * m * n) - 12.0) / 2.5);
  feat[7] = (float)((log2(2.0 * m * n * k) - 15.0) / 3.0);
  memset(feat + MLP_NSHAPE, 0, sizeof(float) * (size_t)mlp_nenc());
  if (NULL != config) { /* the vector kernel has no compiled kt; flush 0 launches the atomic flush */
    const int value[MLP_NPARAM] = {config->variant, config->tm, config->tn, config->nt,
      0 == config->variant ? 0 : config->kt, config->bs, 0 == config->flush ? 2 : config->flush};
    for (h = 0; h < MLP_NPARAM; ++h) {
      const int c = mlp_class(h, value[h]);
      if (0 > c) return EXIT_FAILURE;
      feat[off + c] = 1.f;
      off += mlp_nvals[h];
    }
  }
  return EXIT_SUCCESS;
}

static uint64_t mlp_mix(uint64_t* s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void mlp_free_device(libsmm_acc_mlp_t* m) {
  if (NULL != m->d_params) (void)hipFree(m->d_params);
  if (NULL != m->d_grad) (void)hipFree(m->d_grad);
  if (NULL != m->d_m) (void)hipFree(m->d_m);
  if (NULL != m->d_v) (void)hipFree(m->d_v);
  m->d_params = m->d_grad = m->d_m = m->d_v = NULL;
}

static libsmm_acc_mlp_t* mlp_alloc(int nnets, int nlayers, const int* width) {
  libsmm_acc_mlp_t* const m = (libsmm_acc_mlp_t*)calloc(1, sizeof(libsmm_acc_mlp_t));
  int l;
  if (NULL == m) return NULL;
  m->nnets = nnets;
  m->nlayers = nlayers;
  for (l = 0; l <= nlayers; ++l) m->width[l] = width[l];
  for (l = 0; l < nlayers; ++l) m->nparams += (width[l] + 1) * width[l + 1];
  m->params = (float*)calloc((size_t)nnets * m->nparams, sizeof(float));
  if (NULL == m->params) {
    free(m);
    return NULL;
  }
  return m;
}

LIBXSTREAM_API int libsmm_acc_mlp_create(libsmm_acc_mlp_t** model, int nnets, int hidden, int nhidden,
  unsigned long long seed) {
  int width[MLP_MAXL + 1], l, net;
  libsmm_acc_mlp_t* m;
  uint64_t st = seed;
  if (NULL == model || 0 >= nnets || MLP_MAXNETS < nnets || 0 >= hidden || 64 < hidden || 0 > nhidden ||
      MLP_MAXL <= nhidden) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_create: invalid argument");
  }
  width[0] = MLP_NSHAPE + mlp_nenc();
  for (l = 1; l <= nhidden; ++l) width[l] = hidden;
  width[nhidden + 1] = MLP_NOUT;
  m = mlp_alloc(nnets, nhidden + 1, width);
  if (NULL == m) return EXIT_FAILURE;
  for (net = 0; net < nnets; ++net) { /* He-uniform weights, zero biases; every member its own draw */
    float* const p = m->params + (size_t)net * m->nparams;
    int off = 0;
    for (l = 0; l < m->nlayers; ++l) {
      const int nin = m->width[l], nout = m->width[l + 1];
      const double lim = sqrt(6.0 / nin);
      int i;
      for (i = 0; i < nin * nout; ++i) {
        const double u = (double)(mlp_mix(&st) >> 11) * (1.0 / 9007199254740992.0);
        p[off + i] = (float)((2.0 * u - 1.0) * lim);
      }
      off += (nin + 1) * nout;
    }
  }
  *model = m;
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libsmm_acc_mlp_destroy(libsmm_acc_mlp_t* model) {
  if (NULL != model) {
    pthread_mutex_lock(&lxs_learned_lock);
    if (lxs_learned == model) lxs_learned = NULL;
    if (lxs_learned_owned == model) lxs_learned_owned = NULL;
    pthread_mutex_unlock(&lxs_learned_lock);
    mlp_free_device(model);
    free(model->params);
    free(model);
  }
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libsmm_acc_mlp_info(const libsmm_acc_mlp_t* model, int* nnets, int* nlayers, int* width, int cap) {
  int l;
  if (NULL == model) return -1;
  if (NULL != nnets) *nnets = model->nnets;
  if (NULL != nlayers) *nlayers = model->nlayers;
  for (l = 0; NULL != width && l <= model->nlayers && l < cap; ++l) width[l] = model->width[l];
  return model->nparams;
}

LIBXSTREAM_API int libsmm_acc_mlp_params(libsmm_acc_mlp_t* model, int net, float* params, int cap) {
  if (NULL == model || 0 > net || model->nnets <= net) return -1;
  if (NULL != params) {
    memcpy(params, model->params + (size_t)net * model->nparams,
      sizeof(float) * (size_t)(cap < model->nparams ? cap : model->nparams));
  }
  return model->nparams;
}

LIBXSTREAM_API int libsmm_acc_mlp_set_params(libsmm_acc_mlp_t* model, int net, const float* params, int n) {
  if (NULL == model || 0 > net || model->nnets <= net || NULL == params || n != model->nparams) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_set_params: size mismatch");
  }
  memcpy(model->params + (size_t)net * model->nparams, params, sizeof(float) * (size_t)n);
  model->step[net] = 0;
  if (NULL != model->d_m) { /* the optimizer state of this member restarts */
    (void)hipMemset(model->d_m + (size_t)net * n, 0, sizeof(float) * (size_t)n);
    (void)hipMemset(model->d_v + (size_t)net * n, 0, sizeof(float) * (size_t)n);
  }
  return EXIT_SUCCESS;
}

/* device buffers of one batch: activations per layer, gradients, targets, loss */
typedef struct mlp_batch_t {
  float* act[MLP_MAXL + 1]; /* act[0] = inputs */
  float* dact[MLP_MAXL + 1];
  float *tgt, *loss, *hw, *scratch;
  int batch;
} mlp_batch_t;

static void mlp_batch_free(mlp_batch_t* b) {
  int l;
  for (l = 0; l <= MLP_MAXL; ++l) {
    if (NULL != b->act[l]) (void)hipFree(b->act[l]);
    if (NULL != b->dact[l]) (void)hipFree(b->dact[l]);
  }
  if (NULL != b->tgt) (void)hipFree(b->tgt);
  if (NULL != b->loss) (void)hipFree(b->loss);
  if (NULL != b->hw) (void)hipFree(b->hw);
  if (NULL != b->scratch) (void)hipFree(b->scratch);
  memset(b, 0, sizeof(*b));
}

static int mlp_batch_init(const libsmm_acc_mlp_t* m, mlp_batch_t* b, const float* x, const float* tgt,
  const float* out_weights, int batch, hipStream_t hs) {
  const int nin = m->width[0], nout = m->width[m->nlayers];
  int l, result = 0;
  memset(b, 0, sizeof(*b));
  b->batch = batch;
  for (l = 0; l <= m->nlayers && 0 == result; ++l) {
    result |= lxs_hip(hipMalloc((void**)&b->act[l], sizeof(float) * (size_t)batch * m->width[l]), "mlp: activations");
    if (0 < l) {
      result |= lxs_hip(hipMalloc((void**)&b->dact[l], sizeof(float) * (size_t)batch * m->width[l]), "mlp: gradients");
    }
  }
  if (0 == result) {
    result |= lxs_hip(hipMemcpyAsync(b->act[0], x, sizeof(float) * (size_t)batch * nin, hipMemcpyHostToDevice, hs),
      "mlp: inputs");
  }
  if (0 == result && NULL != tgt) {
    float hw[MLP_NOUT];
    int c;
    for (c = 0; c < nout && c < MLP_NOUT; ++c) hw[c] = (NULL != out_weights ? out_weights[c] : 1.f);
    result |= lxs_hip(hipMalloc((void**)&b->tgt, sizeof(float) * (size_t)batch * nout), "mlp: targets");
    result |= lxs_hip(hipMalloc((void**)&b->loss, sizeof(float) * (size_t)batch), "mlp: loss");
    result |= lxs_hip(hipMalloc((void**)&b->hw, sizeof(float) * (size_t)nout), "mlp: output weights");
    { /* partial weight gradients: sized for the widest layer */
      long most = 0;
      for (l = 0; l < m->nlayers; ++l) {
        const long need = lxs_mlp_scratch(batch, m->width[l], m->width[l + 1]);
        most = (need > most ? need : most);
      }
      result |= lxs_hip(hipMalloc((void**)&b->scratch, sizeof(float) * (size_t)most), "mlp: scratch");
    }
    if (0 == result) {
      result |= lxs_hip(hipMemcpyAsync(b->tgt, tgt, sizeof(float) * (size_t)batch * nout, hipMemcpyHostToDevice, hs), "mlp");
      result |= lxs_hip(hipMemcpyAsync(b->hw, hw, sizeof(float) * (size_t)nout, hipMemcpyHostToDevice, hs), "mlp");
    }
  }
  if (0 == result) result = lxs_hip(hipStreamSynchronize(hs), "mlp"); /* host arrays may go away */
  if (0 != result) mlp_batch_free(b);
  return result;
}

static int mlp_device(libsmm_acc_mlp_t* m, int train, hipStream_t hs) {
  const size_t nb = sizeof(float) * (size_t)m->nnets * m->nparams;
  int result = 0;
  if (NULL == m->d_params) result |= lxs_hip(hipMalloc((void**)&m->d_params, nb), "mlp: parameters");
  if (0 == result) {
    result |= lxs_hip(hipMemcpyAsync(m->d_params, m->params, nb, hipMemcpyHostToDevice, hs), "mlp: parameters");
  }
  if (0 == result && 0 != train && NULL == m->d_grad) {
    result |= lxs_hip(hipMalloc((void**)&m->d_grad, nb), "mlp: gradient");
    result |= lxs_hip(hipMalloc((void**)&m->d_m, nb), "mlp: adam m");
    result |= lxs_hip(hipMalloc((void**)&m->d_v, nb), "mlp: adam v");
    if (0 == result) {
      result |= lxs_hip(hipMemsetAsync(m->d_m, 0, nb, hs), "mlp");
      result |= lxs_hip(hipMemsetAsync(m->d_v, 0, nb, hs), "mlp");
    }
  }
  return result;
}

static int mlp_fwd(const libsmm_acc_mlp_t* m, int net, mlp_batch_t* b, hipStream_t hs) {
  const float* const p = m->d_params + (size_t)net * m->nparams;
  int l, off = 0, result = 0;
  for (l = 0; l < m->nlayers && 0 == result; ++l) {
    result = lxs_mlp_forward(b->act[l], p + off, b->batch, m->width[l], m->width[l + 1], (l + 1 < m->nlayers ? 1 : 0),
      b->act[l + 1], hs);
    off += (m->width[l] + 1) * m->width[l + 1];
  }
  return result;
}

static int mlp_fwd_bwd(const libsmm_acc_mlp_t* m, int net, mlp_batch_t* b, hipStream_t hs) {
  const int L = m->nlayers;
  const float* const p = m->d_params + (size_t)net * m->nparams;
  float* const g = m->d_grad + (size_t)net * m->nparams;
  int l, off = m->nparams, result = mlp_fwd(m, net, b, hs);
  if (0 == result) result = lxs_mlp_loss(b->act[L], b->tgt, b->hw, b->batch, m->width[L], b->dact[L], b->loss, hs);
  for (l = L - 1; l >= 0 && 0 == result; --l) {
    off -= (m->width[l] + 1) * m->width[l + 1];
    /* dact[l+1] holds dL/d(output of layer l); layer l's output is ReLU'd unless it is the last */
    result = lxs_mlp_backward(b->act[l], b->act[l + 1], p + off, b->dact[l + 1], b->batch, m->width[l], m->width[l + 1],
      (l + 1 < L ? 1 : 0), g + off, (0 < l ? b->dact[l] : NULL), b->scratch, hs);
  }
  return result;
}

static double mlp_loss_host(const mlp_batch_t* b, hipStream_t hs) {
  float* const loss = (float*)malloc(sizeof(float) * (size_t)b->batch);
  double sum = NAN;
  int i;
  if (NULL != loss &&
      0 == lxs_hip(hipMemcpyAsync(loss, b->loss, sizeof(float) * (size_t)b->batch, hipMemcpyDeviceToHost, hs), "mlp") &&
      0 == lxs_hip(hipStreamSynchronize(hs), "mlp")) {
    sum = 0;
    for (i = 0; i < b->batch; ++i) sum += loss[i];
    sum /= b->batch;
  }
  free(loss);
  return sum;
}

LIBXSTREAM_API int libsmm_acc_mlp_gradient(libsmm_acc_mlp_t* model, int net, const float* x, const float* targets,
  const float* out_weights, int batch, float* grad, double* loss) {
  hipStream_t hs;
  mlp_batch_t b;
  int result;
  if (NULL == model || 0 > net || model->nnets <= net || NULL == x || NULL == targets || 0 >= batch) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_gradient: invalid argument");
  }
  if (0 != lxs_ensure_device()) return EXIT_FAILURE;
  hs = lxs_stream_hip(NULL);
  result = mlp_device(model, 1, hs);
  if (0 == result) result = mlp_batch_init(model, &b, x, targets, out_weights, batch, hs);
  if (0 == result) {
    result = mlp_fwd_bwd(model, net, &b, hs);
    if (0 == result && NULL != grad) {
      result = lxs_hip(hipMemcpyAsync(grad, model->d_grad + (size_t)net * model->nparams,
        sizeof(float) * (size_t)model->nparams, hipMemcpyDeviceToHost, hs), "mlp: gradient");
    }
    if (0 == result && NULL != loss) *loss = mlp_loss_host(&b, hs);
    if (0 == result) result = lxs_hip(hipStreamSynchronize(hs), "mlp");
    mlp_batch_free(&b);
  }
  return (0 == result ? EXIT_SUCCESS : EXIT_FAILURE);
}

LIBXSTREAM_API int libsmm_acc_mlp_train(libsmm_acc_mlp_t* model, const float* x, const float* targets,
  const float* out_weights, int batch, int epochs, double lr, double weight_decay, double* loss) {
  hipStream_t hs;
  mlp_batch_t b;
  int result, e, net;
  if (NULL == model || NULL == x || NULL == targets || 0 >= batch || 0 > epochs) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_train: invalid argument");
  }
  if (0 != lxs_ensure_device()) return EXIT_FAILURE;
  hs = lxs_stream_hip(NULL);
  result = mlp_device(model, 1, hs);
  if (0 == result) result = mlp_batch_init(model, &b, x, targets, out_weights, batch, hs);
  if (0 == result) {
    double total = 0;
    for (net = 0; net < model->nnets && 0 == result; ++net) {
      const size_t off = (size_t)net * model->nparams;
      for (e = 0; e < epochs && 0 == result; ++e) { /* full-batch Adam steps */
        result = mlp_fwd_bwd(model, net, &b, hs);
        if (0 == result) {
          ++model->step[net];
          result = lxs_mlp_adam(model->d_params + off, model->d_grad + off, model->d_m + off, model->d_v + off,
            model->nparams, (float)lr, 0.9f, 0.999f, 1e-8f, model->step[net], (float)weight_decay, hs);
        }
      }
      if (0 == result && NULL != loss) { /* loss of the member's final parameters */
        result = mlp_fwd_bwd(model, net, &b, hs);
        total += mlp_loss_host(&b, hs);
      }
    }
    if (NULL != loss) *loss = total / model->nnets;
    if (0 == result) {
      result = lxs_hip(hipMemcpyAsync(model->params, model->d_params,
        sizeof(float) * (size_t)model->nnets * model->nparams, hipMemcpyDeviceToHost, hs), "mlp: parameters");
    }
    if (0 == result) result = lxs_hip(hipStreamSynchronize(hs), "mlp");
    mlp_batch_free(&b);
  }
  return (0 == result ? EXIT_SUCCESS : EXIT_FAILURE);
}

LIBXSTREAM_API int libsmm_acc_mlp_forward(libsmm_acc_mlp_t* model, const float* x, int batch, float* out) {
  hipStream_t hs;
  mlp_batch_t b;
  int result;
  if (NULL == model || NULL == x || NULL == out || 0 >= batch) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_forward: invalid argument");
  }
  if (0 != lxs_ensure_device()) return EXIT_FAILURE;
  hs = lxs_stream_hip(NULL);
  result = mlp_device(model, 0, hs);
  if (0 == result) result = mlp_batch_init(model, &b, x, NULL, NULL, batch, hs);
  if (0 == result) {
    const int nout = model->width[model->nlayers];
    const size_t n = (size_t)batch * nout;
    float* const tmp = (float*)malloc(sizeof(float) * n);
    int net;
    size_t i;
    if (NULL == tmp) result = EXIT_FAILURE;
    for (i = 0; i < n && 0 == result; ++i) out[i] = 0.f;
    for (net = 0; net < model->nnets && 0 == result; ++net) { /* ensemble mean */
      result = mlp_fwd(model, net, &b, hs);
      if (0 == result) {
        result = lxs_hip(hipMemcpyAsync(tmp, b.act[model->nlayers], sizeof(float) * n, hipMemcpyDeviceToHost, hs),
          "mlp: output");
      }
      if (0 == result) result = lxs_hip(hipStreamSynchronize(hs), "mlp");
      for (i = 0; i < n && 0 == result; ++i) out[i] += tmp[i] / (float)model->nnets;
    }
    free(tmp);
    mlp_batch_free(&b);
  }
  return (0 == result ? EXIT_SUCCESS : EXIT_FAILURE);
}

LIBXSTREAM_API int libsmm_acc_mlp_save(const libsmm_acc_mlp_t* model, const char* filename) {
  FILE* f;
  int l;
  size_t i;
  if (NULL == model || NULL == filename || NULL == (f = fopen(filename, "w"))) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_save: cannot write file");
  }
  fprintf(f, "LIBSMM_MLP 2 %d %d", model->nnets, model->nlayers);
  for (l = 0; l <= model->nlayers; ++l) fprintf(f, " %d", model->width[l]);
  fprintf(f, "\n");
  for (i = 0; i < (size_t)model->nnets * model->nparams; ++i) fprintf(f, "%.9g\n", (double)model->params[i]);
  return (0 == fclose(f) ? EXIT_SUCCESS : EXIT_FAILURE);
}

LIBXSTREAM_API int libsmm_acc_mlp_load(libsmm_acc_mlp_t** model, const char* filename) {
  FILE* f;
  int version = 0, nnets = 0, nlayers = 0, width[MLP_MAXL + 1], l, ok;
  size_t i;
  libsmm_acc_mlp_t* m = NULL;
  if (NULL == model || NULL == filename || NULL == (f = fopen(filename, "r"))) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_load: cannot read file");
  }
  ok = (3 == fscanf(f, "LIBSMM_MLP %d %d %d", &version, &nnets, &nlayers) && 2 == version && 0 < nnets &&
        MLP_MAXNETS >= nnets && 0 < nlayers && MLP_MAXL >= nlayers);
  for (l = 0; ok && l <= nlayers; ++l) ok = (1 == fscanf(f, "%d", &width[l]) && 0 < width[l] && 4096 > width[l]);
  ok = ok && (MLP_NSHAPE + mlp_nenc()) == width[0] && MLP_NOUT == width[nlayers];
  if (ok) ok = (NULL != (m = mlp_alloc(nnets, nlayers, width)));
  for (i = 0; ok && i < (size_t)nnets * m->nparams; ++i) ok = (1 == fscanf(f, "%f", &m->params[i]));
  fclose(f);
  if (!ok) {
    libsmm_acc_mlp_destroy(m);
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_load: malformed model file");
  }
  *model = m;
  return EXIT_SUCCESS;
}

/*
 * Ensemble mean of the network outputs on the host (out: batch x outputs), from model->params. It runs
 * the same arithmetic as the device forward pass (mlp.hip, mlp_forward): per output, the bias, then one
 * fused multiply-add per input in input order, ReLU on every layer but the last, and the members'
 * outputs added in member order after division by the member count, so both give the same values.
 * The predictor slot of libsmm_acc_process uses it: no device memory, copy or synchronisation on the
 * launch path. Weights are walked input-major (a transposed copy), so that the compiler vectorises
 * across outputs without changing any output's summation order.
 */
#if defined(__SANITIZE_THREAD__) /* a clone's resolver runs before the ThreadSanitizer runtime is up */
#define LXS_MLP_CLONES
#else
#define LXS_MLP_CLONES __attribute__((target_clones("fma", "default")))
#endif
LXS_MLP_CLONES static int mlp_forward_host(const libsmm_acc_mlp_t* model,
  const float* x, int batch, float* out) {
  int wmax = 0, l, net, b, i, o;
  float *wt, *act0, *act1;
  for (l = 0; l <= model->nlayers; ++l) wmax = (wmax < model->width[l] ? model->width[l] : wmax);
  wt = (float*)malloc(sizeof(float) * (size_t)model->nparams);
  act0 = (float*)malloc(sizeof(float) * (size_t)wmax);
  act1 = (float*)malloc(sizeof(float) * (size_t)wmax);
  if (NULL == wt || NULL == act0 || NULL == act1) {
    free(wt);
    free(act0);
    free(act1);
    return EXIT_FAILURE;
  }
  for (i = 0; i < batch * model->width[model->nlayers]; ++i) out[i] = 0.f;
  for (net = 0; net < model->nnets; ++net) {
    const float* const p = model->params + (size_t)net * model->nparams;
    size_t off = 0;
    for (l = 0; l < model->nlayers; ++l) { /* [out][in] weights, then the bias: to [in][out], then the bias */
      const int nin = model->width[l], nout = model->width[l + 1];
      for (o = 0; o < nout; ++o) {
        for (i = 0; i < nin; ++i) wt[off + (size_t)i * nout + o] = p[off + (size_t)o * nin + i];
      }
      for (o = 0; o < nout; ++o) wt[off + (size_t)nin * nout + o] = p[off + (size_t)nout * nin + o];
      off += (size_t)(nin + 1) * nout;
    }
    for (b = 0; b < batch; ++b) {
      const float* in = x + (size_t)b * model->width[0];
      off = 0;
      for (l = 0; l < model->nlayers; ++l) {
        const int nin = model->width[l], nout = model->width[l + 1], relu = (l + 1 < model->nlayers);
        float* const y = (0 == (l & 1) ? act0 : act1);
        const float* const w = wt + off;
        for (o = 0; o < nout; ++o) y[o] = w[(size_t)nin * nout + o];
        for (i = 0; i < nin; ++i) {
          const float xi = in[i];
          const float* const wi = w + (size_t)i * nout;
          for (o = 0; o < nout; ++o) y[o] = fmaf(xi, wi[o], y[o]);
        }
        if (relu) {
          for (o = 0; o < nout; ++o) y[o] = (y[o] < 0.f ? 0.f : y[o]);
        }
        off += (size_t)(nin + 1) * nout;
        in = y;
      }
      {
        const int nout = model->width[model->nlayers];
        for (o = 0; o < nout; ++o) out[(size_t)b * nout + o] += in[o] / (float)model->nnets;
      }
    }
  }
  free(wt);
  free(act0);
  free(act1);
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libsmm_acc_mlp_forward_host(const libsmm_acc_mlp_t* model, const float* x, int batch, float* out) {
  if (NULL == model || NULL == x || NULL == out || 0 >= batch) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_forward_host: invalid argument");
  }
  return mlp_forward_host(model, x, batch, out);
}

/*
 * The ensemble's ranking of every valid configuration of the shape (the set the exhaustive autotuner
 * enumerates): the topk best by predicted relative throughput, best first, with their predicted
 * relative throughput (scores) and predicted GFLOP/s (configs[i].gflops); *count of them.
 */
static int mlp_rank(libsmm_acc_mlp_t* model, libsmm_acc_data_t datatype, int m, int n, int k, int topk,
  libsmm_acc_smm_config_t* configs, double* scores, int* count_out) {
  static const int bss[4] = {16, 32, 64, 128};
  const int nin = MLP_NSHAPE + mlp_nenc();
  int cand[5 * 256], ncand, i, j, b, fl, count = 0, result = EXIT_FAILURE;
  libsmm_acc_smm_config_t* cfgs;
  float *x, *out;
  int* idx;
  *count_out = 0;
  ncand = libsmm_acc_kernel_candidates(datatype, m, n, k, cand, 256);
  if (0 >= ncand || 0 >= topk) return EXIT_FAILURE;
  if (256 < ncand) ncand = 256;
  cfgs = (libsmm_acc_smm_config_t*)malloc(sizeof(libsmm_acc_smm_config_t) * (size_t)ncand * 8);
  x = (float*)malloc(sizeof(float) * (size_t)ncand * 8 * nin);
  out = (float*)malloc(sizeof(float) * (size_t)ncand * 8 * MLP_NOUT);
  idx = (int*)malloc(sizeof(int) * (size_t)ncand * 8);
  if (NULL != cfgs && NULL != x && NULL != out && NULL != idx) {
    for (i = 0; i < ncand; ++i) { /* every valid configuration, as the exhaustive autotuner enumerates them */
      for (b = 0; b < 4; ++b) {
        /* flush 1 launches as the atomic flush unless stacks are declared grouped by C */
        for (fl = (0 != lxs_cfg.grouped ? 1 : 2); fl <= 2; ++fl) {
          libsmm_acc_smm_config_t cfg;
          memset(&cfg, 0, sizeof(cfg));
          cfg.variant = cand[5 * i];
          cfg.tm = cand[5 * i + 1];
          cfg.tn = cand[5 * i + 2];
          cfg.nt = cand[5 * i + 3];
          cfg.kt = cand[5 * i + 4];
          cfg.bs = bss[b];
          cfg.flush = fl;
          if (EXIT_SUCCESS != libsmm_acc_config_valid(datatype, m, n, k, &cfg)) {
            if (0 != cfg.variant) continue;
            for (cfg.kt = (k > 8 ? (k - 1) / 8 * 8 : k - 1); 0 < cfg.kt; cfg.kt -= (cfg.kt > 8 ? 8 : 1)) {
              if (EXIT_SUCCESS == libsmm_acc_config_valid(datatype, m, n, k, &cfg)) break;
            }
            if (0 >= cfg.kt) continue;
          }
          if (EXIT_SUCCESS != libsmm_acc_mlp_features(datatype, m, n, k, &cfg, x + (size_t)count * nin)) continue;
          cfgs[count++] = cfg;
        }
      }
    }
    if (0 < count && EXIT_SUCCESS == mlp_forward_host(model, x, count, out)) {
      const int nsel = (topk < count ? topk : count);
      for (i = 0; i < count; ++i) idx[i] = i;
      for (i = 0; i < nsel; ++i) { /* selection of the nsel best, ties to the earlier candidate */
        int best = i;
        for (j = i + 1; j < count; ++j) {
          if (out[(size_t)idx[j] * MLP_NOUT] > out[(size_t)idx[best] * MLP_NOUT]) best = j;
        }
        j = idx[i];
        idx[i] = idx[best];
        idx[best] = j;
        configs[i] = cfgs[idx[i]];
        configs[i].gflops = exp2((double)out[(size_t)idx[i] * MLP_NOUT + 1]);
        if (NULL != scores) scores[i] = (double)out[(size_t)idx[i] * MLP_NOUT];
      }
      *count_out = nsel;
      result = EXIT_SUCCESS;
    }
  }
  free(cfgs);
  free(x);
  free(out);
  free(idx);
  return result;
}

/* shared by the public predictor call and the registry's predictor slot */
static int mlp_predict(libsmm_acc_mlp_t* model, libsmm_acc_data_t datatype, int m, int n, int k,
  libsmm_acc_smm_config_t* config, double* confidence) {
  double score = 0;
  int count = 0;
  if (EXIT_SUCCESS != mlp_rank(model, datatype, m, n, k, 1, config, &score, &count) || 1 != count) {
    return EXIT_FAILURE;
  }
  if (NULL != confidence) *confidence = (score < 0 ? 0 : (score > 1 ? 1 : score));
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libsmm_acc_mlp_predict_topk(libsmm_acc_mlp_t* model, libsmm_acc_data_t datatype, int m, int n,
  int k, int topk, libsmm_acc_smm_config_t* configs, double* scores, int* count) {
  if (NULL == model || NULL == configs || NULL == count || 0 >= m || 0 >= n || 0 >= k || 0 >= topk) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_predict_topk: invalid argument");
  }
  if (EXIT_SUCCESS != mlp_rank(model, datatype, m, n, k, topk, configs, scores, count)) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_predict_topk: no valid configuration");
  }
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libsmm_acc_mlp_predict_config(libsmm_acc_mlp_t* model, libsmm_acc_data_t datatype, int m, int n, int k,
  libsmm_acc_smm_config_t* config, double* confidence) {
  if (NULL == model || NULL == config || 0 >= m || 0 >= n || 0 >= k) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_predict_config: invalid argument");
  }
  if (EXIT_SUCCESS != mlp_predict(model, datatype, m, n, k, config, confidence)) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_predict_config: no valid configuration");
  }
  return EXIT_SUCCESS;
}

LIBXSTREAM_API int libsmm_acc_learned_set(libsmm_acc_mlp_t* model, double min_confidence) {
  pthread_mutex_lock(&lxs_learned_lock);
  lxs_learned = model;
  lxs_learned_conf = min_confidence;
  pthread_mutex_unlock(&lxs_learned_lock);
  return EXIT_SUCCESS;
}

/* the registry's predictor slot: EXIT_SUCCESS with a configuration when a model is installed and confident */
int lxs_learned_predict(int datatype, int m, int n, int k, libsmm_acc_smm_config_t* config) {
  int result = EXIT_FAILURE;
  pthread_mutex_lock(&lxs_learned_lock);
  if (NULL != lxs_learned) {
    double conf = 0;
    libsmm_acc_smm_config_t cfg;
    if (EXIT_SUCCESS == mlp_predict(lxs_learned, (libsmm_acc_data_t)datatype, m, n, k, &cfg, &conf) &&
        conf >= lxs_learned_conf) {
      *config = cfg;
      result = EXIT_SUCCESS;
    }
    else {
      (void)libxstream_error_consume(NULL, 0);
    }
  }
  pthread_mutex_unlock(&lxs_learned_lock);
  return result;
}

void lxs_learned_load_default(const char* path) {
  libsmm_acc_mlp_t* model = NULL;
  if (EXIT_SUCCESS == libsmm_acc_mlp_load(&model, path)) {
    const char* const env = getenv("LIBXSTREAM_LEARNED_CONF");
    pthread_mutex_lock(&lxs_learned_lock);
    lxs_learned = lxs_learned_owned = model;
    lxs_learned_conf = (NULL != env && '\0' != *env) ? atof(env) : MLP_DEFAULT_CONF;
    pthread_mutex_unlock(&lxs_learned_lock);
  }
  else {
    (void)libxstream_error_consume(NULL, 0);
  }
}

void lxs_learned_finalize(void) {
  libsmm_acc_mlp_t* owned;
  pthread_mutex_lock(&lxs_learned_lock);
  lxs_learned = NULL;
  owned = lxs_learned_owned;
  lxs_learned_owned = NULL;
  pthread_mutex_unlock(&lxs_learned_lock);
  if (NULL != owned) libsmm_acc_mlp_destroy(owned);
}
