/*
 * ORACLE: test infrastructure only. It checks the HIP product path; it is never the product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library.
 *
 * Double-precision restatement of the learned predictor's network and loss: fully connected
 * layers y = act(W x + b) with weights [out][in] followed by the bias [out] per layer, packed layer
 * after layer; ReLU on every layer but the last. The output is cut into head segments: a segment of
 * length 1 is a regression (hw (z - t)^2), a longer one a softmax cross-entropy against a soft
 * target (-hw sum_c t_c log softmax(z)_c). The loss is the batch mean of the per-sample sums.
 *   oracle_mlp_loss_grad  loss and its gradient with respect to every parameter (backpropagation)
 *
 * Pinning (tests/test_oracle_mlp.py): central finite differences of the loss, a one-layer linear
 * network whose gradient has a closed form, and planted mistakes the pins must reject.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MLP_PLANT_NONE 0
#define MLP_PLANT_NORELU 1   /* backward ignores the ReLU mask */
#define MLP_PLANT_NOBIAS 2   /* bias gradient left at zero */

int oracle_mlp_loss_grad(int nlayers, const int* width, const double* params, const double* x, const double* tgt,
  const int* seg, const double* hw, int nseg, int batch, double* grad, double* loss, int plant) {
  double **act = NULL, **dact = NULL, total = 0;
  int l, b, result = 0;
  long off;
  if (0 >= nlayers || 0 >= batch) return 1;
  act = (double**)calloc((size_t)nlayers + 1, sizeof(double*));
  dact = (double**)calloc((size_t)nlayers + 1, sizeof(double*));
  if (NULL == act || NULL == dact) {
    free(act), free(dact);
    return 2;
  }
  for (l = 0; l <= nlayers; ++l) {
    act[l] = (double*)calloc((size_t)batch * width[l], sizeof(double));
    dact[l] = (double*)calloc((size_t)batch * width[l], sizeof(double));
    if (NULL == act[l] || NULL == dact[l]) result = 2;
  }
  if (0 == result) {
    long nparams = 0;
    for (l = 0; l < nlayers; ++l) nparams += (long)(width[l] + 1) * width[l + 1];
    if (NULL != grad) memset(grad, 0, sizeof(double) * (size_t)nparams);
    memcpy(act[0], x, sizeof(double) * (size_t)batch * width[0]);
    /* forward */
    off = 0;
    for (l = 0; l < nlayers; ++l) {
      const int nin = width[l], nout = width[l + 1];
      const double* const w = params + off;
      for (b = 0; b < batch; ++b) {
        int o, i;
        for (o = 0; o < nout; ++o) {
          double acc = w[(long)nout * nin + o];
          for (i = 0; i < nin; ++i) acc += act[l][(long)b * nin + i] * w[(long)o * nin + i];
          act[l + 1][(long)b * nout + o] = (l + 1 < nlayers && acc < 0) ? 0.0 : acc;
        }
      }
      off += (long)(nin + 1) * nout;
    }
    /* loss and its gradient on the output */
    for (b = 0; b < batch; ++b) {
      const int nh = width[nlayers];
      const double* const z = act[nlayers] + (long)b * nh;
      const double* const t = tgt + (long)b * nh;
      double* const dz = dact[nlayers] + (long)b * nh;
      int s;
      for (s = 0; s < nseg; ++s) {
        const int c0 = seg[s], c1 = seg[s + 1];
        int c;
        if (1 == c1 - c0) {
          const double d = z[c0] - t[c0];
          total += hw[s] * d * d;
          dz[c0] = 2 * hw[s] * d / batch;
        }
        else {
          double mx = z[c0], sum = 0, tsum = 0, lse;
          for (c = c0 + 1; c < c1; ++c) mx = (z[c] > mx ? z[c] : mx);
          for (c = c0; c < c1; ++c) sum += exp(z[c] - mx), tsum += t[c];
          lse = mx + log(sum);
          for (c = c0; c < c1; ++c) {
            total -= hw[s] * t[c] * (z[c] - lse);
            dz[c] = hw[s] * (tsum * exp(z[c] - lse) - t[c]) / batch;
          }
        }
      }
    }
    /* backward */
    for (l = nlayers - 1; l >= 0 && NULL != grad; --l) {
      const int nin = width[l], nout = width[l + 1];
      const double* w;
      double* g;
      off -= (long)(nin + 1) * nout;
      w = params + off;
      g = grad + off;
      for (b = 0; b < batch; ++b) {
        int o, i;
        for (o = 0; o < nout; ++o) {
          double d = dact[l + 1][(long)b * nout + o];
          if (l + 1 < nlayers && MLP_PLANT_NORELU != plant && !(act[l + 1][(long)b * nout + o] > 0)) d = 0;
          for (i = 0; i < nin; ++i) {
            g[(long)o * nin + i] += d * act[l][(long)b * nin + i];
            dact[l][(long)b * nin + i] += d * w[(long)o * nin + i];
          }
          if (MLP_PLANT_NOBIAS != plant) g[(long)nout * nin + o] += d;
        }
      }
    }
    if (NULL != loss) *loss = total / batch;
  }
  for (l = 0; l <= nlayers; ++l) {
    free(act[l]);
    free(dact[l]);
  }
  free(act);
  free(dact);
  return result;
}
