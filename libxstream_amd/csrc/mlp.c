/*
 * Learned stack-kernel predictor: an FP32 MLP (8 shape features -> 3 x 64 ReLU -> heads) trained and
 * evaluated with the HIP kernels of mlp.hip.
 *
 * The reference picks parameters for an untuned (m, n, k) from the nearest tuned triplet
 * (samples/smm/smm_kernel.c:116-148, :206-213) or falls back to defaults. Here the network's heads
 * (variant, tm, tn, nt, kt, bs, flush: softmax classifiers; log2 GFLOP/s: regression) score every
 * compiled configuration that is valid for the shape by the sum of its per-head log-probabilities;
 * the best one is taken when its share of the candidates' probability mass reaches the predictor's
 * confidence threshold, otherwise the default heuristic stays in place (and the exhaustive
 * autotuner, tune_multiply, remains the fallback for such shapes).
 */
#include "lxs_internal.h"
#include <math.h>
#include <stdint.h>
#include <string.h>

#define MLP_NFEAT 8
#define MLP_MAXL 8

/* heads: (name, class values); a head with no values is the log2 GFLOP/s regression */
static const int mlp_variant[] = {0, 1};
static const int mlp_tm[] = {1, 2, 3, 4, 5};
static const int mlp_tn[] = {1, 2, 3, 4, 5};
static const int mlp_nt[] = {64, 128, 192, 256, 320, 512};
static const int mlp_kt[] = {0, 8, 12, 16, 24};
static const int mlp_bs[] = {16, 32, 64, 128};
static const int mlp_flush[] = {1, 2};
#define MLP_NHEADS 8
static const int* const mlp_vals[MLP_NHEADS] = {mlp_variant, mlp_tm, mlp_tn, mlp_nt, mlp_kt, mlp_bs, mlp_flush, NULL};
static const int mlp_nvals[MLP_NHEADS] = {2, 5, 5, 6, 5, 4, 2, 1};

struct libsmm_acc_mlp_t {
  int nlayers;          /* weight layers */
  int width[MLP_MAXL + 1]; /* width[0] = features, width[nlayers] = outputs */
  int nparams;
  float* params;        /* host copy (authoritative between calls) */
  /* device state for training/inference */
  float *d_params, *d_grad, *d_m, *d_v;
  int step;
};

static libsmm_acc_mlp_t* lxs_learned = NULL;
static libsmm_acc_mlp_t* lxs_learned_owned = NULL; /* the shipped model, loaded at init */
static double lxs_learned_conf = 0.0;
#define MLP_DEFAULT_CONF 0.0 /* confidence threshold of the shipped model (LIBXSTREAM_LEARNED_CONF) */
static pthread_mutex_t lxs_learned_lock = PTHREAD_MUTEX_INITIALIZER;

static int mlp_nout(void) {
  int h, n = 0;
  for (h = 0; h < MLP_NHEADS; ++h) n += mlp_nvals[h];
  return n;
}

LIBXSTREAM_API int libsmm_acc_mlp_heads(int* seg, int* values, int cap) {
  int h, n = 0, c = 0;
  for (h = 0; h < MLP_NHEADS; ++h) {
    if (NULL != seg && h < cap) seg[h] = n;
    if (NULL != values) {
      int v;
      for (v = 0; v < mlp_nvals[h]; ++v) {
        if (c < cap) values[c] = (NULL != mlp_vals[h] ? mlp_vals[h][v] : -1);
        ++c;
      }
    }
    n += mlp_nvals[h];
  }
  if (NULL != seg && MLP_NHEADS < cap) seg[MLP_NHEADS] = n;
  return MLP_NHEADS;
}

LIBXSTREAM_API int libsmm_acc_mlp_features(libsmm_acc_data_t datatype, int m, int n, int k, float* feat) {
  const double ts = (dbcsr_type_real_8 == datatype ? 8.0 : (dbcsr_type_real_4 == datatype ? 4.0 : 0.0));
  if (NULL == feat || 0 >= m || 0 >= n || 0 >= k || 0 == ts) return EXIT_FAILURE;
  feat[0] = (float)((log2((double)m) - 4.0) / 1.5);
  feat[1] = (float)((log2((double)n) - 4.0) / 1.5);
  feat[2] = (float)((log2((double)k) - 4.0) / 1.5);
  feat[3] = (8.0 == ts ? 1.f : -1.f);
  feat[4] = (float)((log2(ts * m * k) - 12.0) / 2.5);
  feat[5] = (float)((log2(ts * k * n) - 12.0) / 2.5);
  feat[6] = (float)((log2(ts * m * n) - 12.0) / 2.5);
  feat[7] = (float)((log2(2.0 * m * n * k) - 15.0) / 3.0);
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

LIBXSTREAM_API int libsmm_acc_mlp_create(libsmm_acc_mlp_t** model, int hidden, int nhidden, unsigned long long seed) {
  libsmm_acc_mlp_t* m;
  int l, off = 0;
  uint64_t st = seed;
  if (NULL == model || 0 >= hidden || 0 > nhidden || MLP_MAXL <= nhidden) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_create: invalid argument");
  }
  m = (libsmm_acc_mlp_t*)calloc(1, sizeof(*m));
  if (NULL == m) return EXIT_FAILURE;
  m->nlayers = nhidden + 1;
  m->width[0] = MLP_NFEAT;
  for (l = 1; l <= nhidden; ++l) m->width[l] = hidden;
  m->width[m->nlayers] = mlp_nout();
  for (l = 0; l < m->nlayers; ++l) m->nparams += (m->width[l] + 1) * m->width[l + 1];
  m->params = (float*)calloc((size_t)m->nparams, sizeof(float));
  if (NULL == m->params) {
    free(m);
    return EXIT_FAILURE;
  }
  for (l = 0; l < m->nlayers; ++l) { /* He-uniform weights, zero biases */
    const int nin = m->width[l], nout = m->width[l + 1];
    const double lim = sqrt(6.0 / nin);
    int i;
    for (i = 0; i < nin * nout; ++i) {
      const double u = (double)(mlp_mix(&st) >> 11) * (1.0 / 9007199254740992.0);
      m->params[off + i] = (float)((2.0 * u - 1.0) * lim);
    }
    off += (nin + 1) * nout;
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

LIBXSTREAM_API int libsmm_acc_mlp_params(libsmm_acc_mlp_t* model, float* params, int cap) {
  if (NULL == model) return -1;
  if (NULL != params) memcpy(params, model->params, sizeof(float) * (size_t)(cap < model->nparams ? cap : model->nparams));
  return model->nparams;
}

LIBXSTREAM_API int libsmm_acc_mlp_set_params(libsmm_acc_mlp_t* model, const float* params, int n) {
  if (NULL == model || NULL == params || n != model->nparams) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_set_params: size mismatch");
  }
  memcpy(model->params, params, sizeof(float) * (size_t)n);
  model->step = 0;
  if (NULL != model->d_m) { /* optimizer state restarts with new parameters */
    (void)hipMemset(model->d_m, 0, sizeof(float) * (size_t)n);
    (void)hipMemset(model->d_v, 0, sizeof(float) * (size_t)n);
  }
  return EXIT_SUCCESS;
}

/* device buffers of one batch: activations per layer, gradients, targets, loss */
typedef struct mlp_batch_t {
  float* act[MLP_MAXL + 1]; /* act[0] = features */
  float* dact[MLP_MAXL + 1];
  float *tgt, *loss, *hw;
  int* seg;
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
  if (NULL != b->seg) (void)hipFree(b->seg);
  memset(b, 0, sizeof(*b));
}

static int mlp_batch_init(const libsmm_acc_mlp_t* m, mlp_batch_t* b, const float* x, const float* tgt,
  const float* head_weights, int batch, hipStream_t hs) {
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
    result |= lxs_hip(hipMemcpyAsync(b->act[0], x, sizeof(float) * (size_t)batch * MLP_NFEAT, hipMemcpyHostToDevice, hs),
      "mlp: features");
  }
  if (0 == result && NULL != tgt) {
    int seg[MLP_NHEADS + 1], h;
    float hw[MLP_NHEADS];
    const int nout = m->width[m->nlayers];
    libsmm_acc_mlp_heads(seg, NULL, MLP_NHEADS + 1);
    for (h = 0; h < MLP_NHEADS; ++h) hw[h] = (NULL != head_weights ? head_weights[h] : 1.f);
    result |= lxs_hip(hipMalloc((void**)&b->tgt, sizeof(float) * (size_t)batch * nout), "mlp: targets");
    result |= lxs_hip(hipMalloc((void**)&b->loss, sizeof(float) * (size_t)batch), "mlp: loss");
    result |= lxs_hip(hipMalloc((void**)&b->hw, sizeof(hw)), "mlp: head weights");
    result |= lxs_hip(hipMalloc((void**)&b->seg, sizeof(seg)), "mlp: segments");
    if (0 == result) {
      result |= lxs_hip(hipMemcpyAsync(b->tgt, tgt, sizeof(float) * (size_t)batch * nout, hipMemcpyHostToDevice, hs), "mlp");
      result |= lxs_hip(hipMemcpyAsync(b->hw, hw, sizeof(hw), hipMemcpyHostToDevice, hs), "mlp");
      result |= lxs_hip(hipMemcpyAsync(b->seg, seg, sizeof(seg), hipMemcpyHostToDevice, hs), "mlp");
      result |= lxs_hip(hipStreamSynchronize(hs), "mlp");
    }
  }
  if (0 != result) mlp_batch_free(b);
  return result;
}

static int mlp_device(libsmm_acc_mlp_t* m, int train, hipStream_t hs) {
  const size_t nb = sizeof(float) * (size_t)m->nparams;
  int result = 0;
  if (NULL == m->d_params) {
    result |= lxs_hip(hipMalloc((void**)&m->d_params, nb), "mlp: parameters");
  }
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

static int mlp_fwd(const libsmm_acc_mlp_t* m, mlp_batch_t* b, hipStream_t hs) {
  int l, off = 0, result = 0;
  for (l = 0; l < m->nlayers && 0 == result; ++l) {
    result = lxs_mlp_forward(b->act[l], m->d_params + off, b->batch, m->width[l], m->width[l + 1],
      (l + 1 < m->nlayers ? 1 : 0), b->act[l + 1], hs);
    off += (m->width[l] + 1) * m->width[l + 1];
  }
  return result;
}

static int mlp_fwd_bwd(const libsmm_acc_mlp_t* m, mlp_batch_t* b, hipStream_t hs) {
  int seg_n = MLP_NHEADS, l, off = m->nparams, result = mlp_fwd(m, b, hs);
  const int L = m->nlayers;
  if (0 == result) {
    result = lxs_mlp_loss(b->act[L], b->tgt, b->seg, b->hw, seg_n, b->batch, m->width[L], b->dact[L], b->loss, hs);
  }
  for (l = L - 1; l >= 0 && 0 == result; --l) {
    off -= (m->width[l] + 1) * m->width[l + 1];
    /* dact[l+1] holds dL/d(output of layer l); layer l's output is ReLU'd unless it is the last */
    result = lxs_mlp_backward(b->act[l], b->act[l + 1], m->d_params + off, b->dact[l + 1], b->batch, m->width[l],
      m->width[l + 1], (l + 1 < L ? 1 : 0), m->d_grad + off, (0 < l ? b->dact[l] : NULL), hs);
  }
  return result;
}

static double mlp_loss_host(const mlp_batch_t* b, hipStream_t hs) {
  float* const loss = (float*)malloc(sizeof(float) * (size_t)b->batch);
  double sum = NAN;
  int i;
  if (NULL != loss && 0 == lxs_hip(hipMemcpyAsync(loss, b->loss, sizeof(float) * (size_t)b->batch, hipMemcpyDeviceToHost, hs), "mlp") &&
      0 == lxs_hip(hipStreamSynchronize(hs), "mlp")) {
    sum = 0;
    for (i = 0; i < b->batch; ++i) sum += loss[i];
    sum /= b->batch;
  }
  free(loss);
  return sum;
}

LIBXSTREAM_API int libsmm_acc_mlp_gradient(libsmm_acc_mlp_t* model, const float* x, const float* targets,
  const float* head_weights, int batch, float* grad, double* loss) {
  hipStream_t hs;
  mlp_batch_t b;
  int result;
  if (NULL == model || NULL == x || NULL == targets || 0 >= batch) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_gradient: invalid argument");
  }
  if (0 != lxs_ensure_device()) return EXIT_FAILURE;
  hs = lxs_stream_hip(NULL);
  result = mlp_device(model, 1, hs);
  if (0 == result) result = mlp_batch_init(model, &b, x, targets, head_weights, batch, hs);
  if (0 == result) {
    result = mlp_fwd_bwd(model, &b, hs);
    if (0 == result && NULL != grad) {
      result = lxs_hip(hipMemcpyAsync(grad, model->d_grad, sizeof(float) * (size_t)model->nparams,
        hipMemcpyDeviceToHost, hs), "mlp: gradient");
    }
    if (0 == result && NULL != loss) *loss = mlp_loss_host(&b, hs);
    if (0 == result) result = lxs_hip(hipStreamSynchronize(hs), "mlp");
    mlp_batch_free(&b);
  }
  return (0 == result ? EXIT_SUCCESS : EXIT_FAILURE);
}

LIBXSTREAM_API int libsmm_acc_mlp_train(libsmm_acc_mlp_t* model, const float* x, const float* targets,
  const float* head_weights, int batch, int epochs, double lr, double weight_decay, double* loss) {
  hipStream_t hs;
  mlp_batch_t b;
  int result, e;
  if (NULL == model || NULL == x || NULL == targets || 0 >= batch || 0 > epochs) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_train: invalid argument");
  }
  if (0 != lxs_ensure_device()) return EXIT_FAILURE;
  hs = lxs_stream_hip(NULL);
  result = mlp_device(model, 1, hs);
  if (0 == result) result = mlp_batch_init(model, &b, x, targets, head_weights, batch, hs);
  if (0 == result) {
    for (e = 0; e < epochs && 0 == result; ++e) { /* full-batch Adam steps */
      result = mlp_fwd_bwd(model, &b, hs);
      if (0 == result) {
        ++model->step;
        result = lxs_mlp_adam(model->d_params, model->d_grad, model->d_m, model->d_v, model->nparams, (float)lr, 0.9f,
          0.999f, 1e-8f, model->step, (float)weight_decay, hs);
      }
    }
    if (0 == result && NULL != loss) { /* loss of the final parameters */
      result = mlp_fwd_bwd(model, &b, hs);
      *loss = mlp_loss_host(&b, hs);
    }
    if (0 == result) {
      result = lxs_hip(hipMemcpyAsync(model->params, model->d_params, sizeof(float) * (size_t)model->nparams,
        hipMemcpyDeviceToHost, hs), "mlp: parameters");
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
    result = mlp_fwd(model, &b, hs);
    if (0 == result) {
      result = lxs_hip(hipMemcpyAsync(out, b.act[model->nlayers], sizeof(float) * (size_t)batch * model->width[model->nlayers],
        hipMemcpyDeviceToHost, hs), "mlp: output");
    }
    if (0 == result) result = lxs_hip(hipStreamSynchronize(hs), "mlp");
    mlp_batch_free(&b);
  }
  return (0 == result ? EXIT_SUCCESS : EXIT_FAILURE);
}

LIBXSTREAM_API int libsmm_acc_mlp_save(const libsmm_acc_mlp_t* model, const char* filename) {
  FILE* f;
  int l, i;
  if (NULL == model || NULL == filename || NULL == (f = fopen(filename, "w"))) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_save: cannot write file");
  }
  fprintf(f, "LIBSMM_MLP 1 %d", model->nlayers);
  for (l = 0; l <= model->nlayers; ++l) fprintf(f, " %d", model->width[l]);
  fprintf(f, "\n");
  for (i = 0; i < model->nparams; ++i) fprintf(f, "%.9g\n", (double)model->params[i]);
  return (0 == fclose(f) ? EXIT_SUCCESS : EXIT_FAILURE);
}

LIBXSTREAM_API int libsmm_acc_mlp_load(libsmm_acc_mlp_t** model, const char* filename) {
  FILE* f;
  int version = 0, nlayers = 0, width[MLP_MAXL + 1], l, i, ok = 1;
  libsmm_acc_mlp_t* m = NULL;
  if (NULL == model || NULL == filename || NULL == (f = fopen(filename, "r"))) {
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_load: cannot read file");
  }
  ok = (2 == fscanf(f, "LIBSMM_MLP %d %d", &version, &nlayers) && 1 == version && 0 < nlayers && MLP_MAXL >= nlayers);
  for (l = 0; ok && l <= nlayers; ++l) ok = (1 == fscanf(f, "%d", &width[l]) && 0 < width[l]);
  ok = ok && MLP_NFEAT == width[0] && mlp_nout() == width[nlayers];
  for (l = 2; ok && l < nlayers; ++l) ok = (width[l] == width[1]);
  if (ok) ok = (EXIT_SUCCESS == libsmm_acc_mlp_create(&m, 1 < nlayers ? width[1] : 1, nlayers - 1, 0));
  for (i = 0; ok && i < m->nparams; ++i) ok = (1 == fscanf(f, "%f", &m->params[i]));
  fclose(f);
  if (!ok) {
    libsmm_acc_mlp_destroy(m);
    return lxs_error(EXIT_FAILURE, "libsmm_acc_mlp_load: malformed model file");
  }
  *model = m;
  return EXIT_SUCCESS;
}

static int mlp_class(int head, int value) {
  int v;
  for (v = 0; v < mlp_nvals[head]; ++v) {
    if (mlp_vals[head][v] == value) return v;
  }
  return -1;
}

/* shared by the public predictor call and the registry's predictor slot */
static int mlp_predict(libsmm_acc_mlp_t* model, libsmm_acc_data_t datatype, int m, int n, int k,
  libsmm_acc_smm_config_t* config, double* confidence) {
  static const int bss[4] = {16, 32, 64, 128};
  float feat[MLP_NFEAT], out[64], lsm[64];
  int seg[MLP_NHEADS + 1], cand[5 * 256], ncand, h, i, b, fl, found = 0;
  double best = -INFINITY, total = 0;
  libsmm_acc_smm_config_t pick;
  if (EXIT_SUCCESS != libsmm_acc_mlp_features(datatype, m, n, k, feat)) return EXIT_FAILURE;
  if (64 < mlp_nout() || EXIT_SUCCESS != libsmm_acc_mlp_forward(model, feat, 1, out)) return EXIT_FAILURE;
  libsmm_acc_mlp_heads(seg, NULL, MLP_NHEADS + 1);
  for (h = 0; h < MLP_NHEADS; ++h) { /* log-softmax per classification head */
    const int c0 = seg[h], c1 = seg[h + 1];
    double mx = out[c0], sum = 0;
    if (NULL == mlp_vals[h]) continue;
    for (i = c0 + 1; i < c1; ++i) mx = (out[i] > mx ? out[i] : mx);
    for (i = c0; i < c1; ++i) sum += exp(out[i] - mx);
    for (i = c0; i < c1; ++i) lsm[i] = (float)(out[i] - mx - log(sum));
  }
  ncand = libsmm_acc_kernel_candidates(datatype, m, n, k, cand, 256);
  memset(&pick, 0, sizeof(pick));
  for (i = 0; i < ncand && i < 256; ++i) {
    for (b = 0; b < 4; ++b) {
      for (fl = 1; fl <= 2; ++fl) {
        const int cls[7] = {mlp_class(0, cand[5 * i]), mlp_class(1, cand[5 * i + 1]), mlp_class(2, cand[5 * i + 2]),
          mlp_class(3, cand[5 * i + 3]), mlp_class(4, cand[5 * i + 4]), mlp_class(5, bss[b]), mlp_class(6, fl)};
        libsmm_acc_smm_config_t cfg;
        double score = 0;
        for (h = 0; h < 7; ++h) {
          if (0 > cls[h]) break;
          score += lsm[seg[h] + cls[h]];
        }
        if (7 > h) continue;
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
        total += exp(score);
        if (score > best) {
          best = score;
          pick = cfg;
          found = 1;
        }
      }
    }
  }
  if (!found) return EXIT_FAILURE;
  pick.gflops = exp2((double)out[seg[MLP_NHEADS - 1]]);
  *config = pick;
  if (NULL != confidence) *confidence = exp(best) / total;
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
