/*
 * FP32 multi-layer perceptron for the learned stack-kernel predictor: forward, loss gradient,
 * backward and Adam update as HIP kernels (gfx950). The network is small (8 -> 64 -> 64 -> 64 -> H),
 * so every kernel maps one thread to one output element and loops over the reduction dimension
 * in a fixed order: results are deterministic run to run.
 *
 * Layouts: activations row-major [batch][features]; layer weights row-major [out][in] followed by
 * the bias [out]; all layers of a network are packed in one parameter vector.
 * Heads: the output vector is cut into segments; a segment of length 1 is a regression output
 * (squared error), a longer one a softmax classifier (cross-entropy against soft targets).
 */
#include <hip/hip_runtime.h>
#include "lxs_kernels.h"

namespace {

/* y[b][o] = act(sum_i x[b][i] w[o][i] + bias[o]) */
__global__ void __launch_bounds__(256) mlp_forward(const float* __restrict__ x, const float* __restrict__ w,
  int batch, int nin, int nout, int relu, float* __restrict__ y) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < batch * nout) {
    const int b = t / nout, o = t - b * nout;
    const float* const xr = x + (long)b * nin;
    const float* const wr = w + (long)o * nin;
    float acc = w[(long)nout * nin + o];
    for (int i = 0; i < nin; ++i) acc = fmaf(xr[i], wr[i], acc);
    y[t] = (0 != relu && acc < 0.f) ? 0.f : acc;
  }
}

/* loss gradient on the output z[b][h] against targets tgt[b][h]; per-sample loss to loss[b].
   seg[s] = first column of segment s (nseg + 1 entries), hw[s] = weight of segment s. */
__global__ void __launch_bounds__(256) mlp_loss(const float* __restrict__ z, const float* __restrict__ tgt,
  const int* __restrict__ seg, const float* __restrict__ hw, int nseg, int batch, int nh, float* __restrict__ dz,
  float* __restrict__ loss) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b < batch) {
    const float* const zr = z + (long)b * nh;
    const float* const tr = tgt + (long)b * nh;
    float* const dr = dz + (long)b * nh;
    const float inv = 1.f / (float)batch;
    float l = 0.f;
    for (int s = 0; s < nseg; ++s) {
      const int c0 = seg[s], c1 = seg[s + 1];
      if (1 == c1 - c0) { /* regression */
        const float d = zr[c0] - tr[c0];
        l += hw[s] * d * d;
        dr[c0] = 2.f * hw[s] * d * inv;
      }
      else { /* softmax cross-entropy with soft targets (summing to 1, or all 0 = no label) */
        float mx = zr[c0], sum = 0.f, tsum = 0.f;
        for (int c = c0 + 1; c < c1; ++c) mx = fmaxf(mx, zr[c]);
        for (int c = c0; c < c1; ++c) sum += expf(zr[c] - mx);
        for (int c = c0; c < c1; ++c) tsum += tr[c];
        const float lse = mx + logf(sum);
        for (int c = c0; c < c1; ++c) {
          const float p = expf(zr[c] - mx) / sum;
          l -= hw[s] * tr[c] * (zr[c] - lse);
          dr[c] = hw[s] * (tsum * p - tr[c]) * inv;
        }
      }
    }
    loss[b] = l;
  }
}

/* dz = dy * (y > 0) in place (ReLU layers) */
__global__ void __launch_bounds__(256) mlp_relu_grad(const float* __restrict__ y, float* __restrict__ dy, int n) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < n && !(y[t] > 0.f)) dy[t] = 0.f;
}

/* dw[o][i] = sum_b dz[b][o] x[b][i]; db[o] = sum_b dz[b][o] */
__global__ void __launch_bounds__(256) mlp_wgrad(const float* __restrict__ x, const float* __restrict__ dz, int batch,
  int nin, int nout, float* __restrict__ dw) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < nout * (nin + 1)) {
    float acc = 0.f;
    if (t < nout * nin) {
      const int o = t / nin, i = t - o * nin;
      for (int b = 0; b < batch; ++b) acc = fmaf(dz[(long)b * nout + o], x[(long)b * nin + i], acc);
    }
    else {
      const int o = t - nout * nin;
      for (int b = 0; b < batch; ++b) acc += dz[(long)b * nout + o];
    }
    dw[t] = acc;
  }
}

/* dx[b][i] = sum_o dz[b][o] w[o][i] */
__global__ void __launch_bounds__(256) mlp_xgrad(const float* __restrict__ dz, const float* __restrict__ w, int batch,
  int nin, int nout, float* __restrict__ dx) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < batch * nin) {
    const int b = t / nin, i = t - b * nin;
    float acc = 0.f;
    for (int o = 0; o < nout; ++o) acc = fmaf(dz[(long)b * nout + o], w[(long)o * nin + i], acc);
    dx[t] = acc;
  }
}

/* Adam (Kingma and Ba, 2015) with bias correction; step counts from 1 */
__global__ void __launch_bounds__(256) mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
  float* __restrict__ v, int n, float lr, float beta1, float beta2, float eps, float c1, float c2, float wd) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t < n) {
    const float gt = g[t] + wd * p[t];
    const float mt = beta1 * m[t] + (1.f - beta1) * gt;
    const float vt = beta2 * v[t] + (1.f - beta2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
  }
}

inline unsigned int blocks(long n) { return (unsigned int)((n + 255) / 256); }

} /* namespace */

extern "C" int lxs_mlp_forward(const float* x, const float* w, int batch, int nin, int nout, int relu, float* y,
  void* hip_stream) {
  if (0 >= batch) return 0;
  mlp_forward<<<blocks((long)batch * nout), 256, 0, (hipStream_t)hip_stream>>>(x, w, batch, nin, nout, relu, y);
  return (int)hipGetLastError();
}

extern "C" int lxs_mlp_loss(const float* z, const float* tgt, const int* seg, const float* hw, int nseg, int batch,
  int nh, float* dz, float* loss, void* hip_stream) {
  if (0 >= batch) return 0;
  mlp_loss<<<blocks(batch), 256, 0, (hipStream_t)hip_stream>>>(z, tgt, seg, hw, nseg, batch, nh, dz, loss);
  return (int)hipGetLastError();
}

extern "C" int lxs_mlp_backward(const float* x, const float* y, const float* w, float* dz, int batch, int nin, int nout,
  int relu, float* dw, float* dx, void* hip_stream) {
  const hipStream_t st = (hipStream_t)hip_stream;
  if (0 >= batch) return 0;
  if (0 != relu) mlp_relu_grad<<<blocks((long)batch * nout), 256, 0, st>>>(y, dz, batch * nout);
  mlp_wgrad<<<blocks((long)nout * (nin + 1)), 256, 0, st>>>(x, dz, batch, nin, nout, dw);
  if (NULL != dx) mlp_xgrad<<<blocks((long)batch * nin), 256, 0, st>>>(dz, w, batch, nin, nout, dx);
  return (int)hipGetLastError();
}

extern "C" int lxs_mlp_adam(float* p, const float* g, float* m, float* v, int n, float lr, float beta1, float beta2,
  float eps, int step, float wd, void* hip_stream) {
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  if (0 >= n) return 0;
  mlp_adam<<<blocks(n), 256, 0, (hipStream_t)hip_stream>>>(p, g, m, v, n, lr, beta1, beta2, eps, c1, c2, wd);
  return (int)hipGetLastError();
}
