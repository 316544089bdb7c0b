/*
 * FP32 multi-layer perceptron of the learned stack-kernel predictor: forward, loss gradient,
 * backward and Adam update as HIP kernels (gfx950).
 *
 * Layouts: activations row-major [batch][features]; layer weights row-major [out][in] followed by
 * the bias [out]; the layers of one network are packed in one parameter vector.
 * Every reduction runs in a fixed order (rows in order within a chunk, chunks in order), so a
 * training run gives the same parameters every time.
 */
#include <hip/hip_runtime.h>
#include "lxs_kernels.h"

namespace {

/* y[b][o] = act(sum_i x[b][i] w[o][i] + bias[o]) */
__global__ void __launch_bounds__(256) mlp_forward(const float* __restrict__ x, const float* __restrict__ w,
  int batch, int nin, int nout, int relu, float* __restrict__ y) {
  const long t = blockIdx.x * 256L + threadIdx.x;
  if (t < (long)batch * nout) {
    const int b = (int)(t / nout), o = (int)(t - (long)b * nout);
    const float* const xr = x + (long)b * nin;
    const float* const wr = w + (long)o * nin;
    float acc = w[(long)nout * nin + o];
    for (int i = 0; i < nin; ++i) acc = fmaf(xr[i], wr[i], acc);
    y[t] = (0 != relu && acc < 0.f) ? 0.f : acc;
  }
}

/* weighted squared error per output column: loss[b] = sum_c hw[c] (z - t)^2, dz = 2 hw (z - t) / batch */
__global__ void __launch_bounds__(256) mlp_loss(const float* __restrict__ z, const float* __restrict__ tgt,
  const float* __restrict__ hw, int batch, int nh, float* __restrict__ dz, float* __restrict__ loss) {
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b < batch) {
    const float inv = 1.f / (float)batch;
    float l = 0.f;
    for (int c = 0; c < nh; ++c) {
      const long o = (long)b * nh + c;
      const float d = z[o] - tgt[o];
      l += hw[c] * d * d;
      dz[o] = 2.f * hw[c] * d * inv;
    }
    loss[b] = l;
  }
}

/* dz = dy * (y > 0) in place (ReLU layers) */
__global__ void __launch_bounds__(256) mlp_relu_grad(const float* __restrict__ y, float* __restrict__ dy, long n) {
  const long t = blockIdx.x * 256L + threadIdx.x;
  if (t < n && !(y[t] > 0.f)) dy[t] = 0.f;
}

/* partial weight gradients of one chunk of rows, in the parameter layout (weights [out][in], then the
   bias [out]): part[c][o nin + i] = sum over the chunk of dz[b][o] x[b][i], part[c][nout nin + o] =
   sum of dz[b][o]; rows walked in order, each thread owns a fixed set of outputs */
__global__ void __launch_bounds__(256) mlp_wgrad_part(const float* __restrict__ x, const float* __restrict__ dz,
  int batch, int nin, int nout, int rows_per_chunk, float* __restrict__ part) {
  constexpr int PER = 17; /* outputs per thread: 256 * 17 >= 64 * 65 */
  const int np = nout * (nin + 1), b0 = blockIdx.x * rows_per_chunk;
  const int b1 = (b0 + rows_per_chunk < batch ? b0 + rows_per_chunk : batch);
  float acc[PER];
  int oo[PER], ii[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int p = threadIdx.x + 256 * q;
    const bool bias = (p >= nout * nin);
    oo[q] = (p < np ? (bias ? p - nout * nin : p / nin) : 0);
    ii[q] = (p < np && !bias ? p - (p / nin) * nin : nin); /* nin: the bias (x = 1) */
    acc[q] = 0.f;
  }
  for (int b = b0; b < b1; ++b) {
    const float* const xr = x + (long)b * nin;
    const float* const dr = dz + (long)b * nout;
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = fmaf(dr[oo[q]], (ii[q] < nin ? xr[ii[q]] : 1.f), acc[q]);
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int p = threadIdx.x + 256 * q;
    if (p < np) part[(long)blockIdx.x * np + p] = acc[q];
  }
}

/* dw[p] = sum over chunks of part[c][p], chunks in order */
__global__ void __launch_bounds__(256) mlp_wgrad_sum(const float* __restrict__ part, int nchunks, int np,
  float* __restrict__ dw) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p < np) {
    float acc = 0.f;
    for (int c = 0; c < nchunks; ++c) acc += part[(long)c * np + p];
    dw[p] = acc;
  }
}

/* dx[b][i] = sum_o dz[b][o] w[o][i] */
__global__ void __launch_bounds__(256) mlp_xgrad(const float* __restrict__ dz, const float* __restrict__ w, int batch,
  int nin, int nout, float* __restrict__ dx) {
  const long t = blockIdx.x * 256L + threadIdx.x;
  if (t < (long)batch * nin) {
    const int b = (int)(t / nin), i = (int)(t - (long)b * nin);
    float acc = 0.f;
    for (int o = 0; o < nout; ++o) acc = fmaf(dz[(long)b * nout + o], w[(long)o * nin + i], acc);
    dx[t] = acc;
  }
}

/* Adam (Kingma and Ba, 2015) with bias correction and L2 weight decay; step counts from 1 */
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

extern "C" int lxs_mlp_loss(const float* z, const float* tgt, const float* hw, int batch, int nh, float* dz,
  float* loss, void* hip_stream) {
  if (0 >= batch) return 0;
  mlp_loss<<<blocks(batch), 256, 0, (hipStream_t)hip_stream>>>(z, tgt, hw, batch, nh, dz, loss);
  return (int)hipGetLastError();
}

/* scratch floats lxs_mlp_backward needs for its partial weight gradients */
extern "C" long lxs_mlp_scratch(int batch, int nin, int nout) {
  const int nchunks = (batch + 399) / 400 < 256 ? (batch + 399) / 400 : 256;
  return (long)(nchunks > 0 ? nchunks : 1) * nout * (nin + 1);
}

extern "C" int lxs_mlp_backward(const float* x, const float* y, const float* w, float* dz, int batch, int nin, int nout,
  int relu, float* dw, float* dx, float* scratch, void* hip_stream) {
  const hipStream_t st = (hipStream_t)hip_stream;
  const int np = nout * (nin + 1);
  const int nchunks = (int)(lxs_mlp_scratch(batch, nin, nout) / np), rows = (batch + nchunks - 1) / nchunks;
  if (0 >= batch) return 0;
  if (64 * 65 < np) return LXS_ELIMIT;
  if (0 != relu) mlp_relu_grad<<<blocks((long)batch * nout), 256, 0, st>>>(y, dz, (long)batch * nout);
  mlp_wgrad_part<<<(unsigned int)nchunks, 256, 0, st>>>(x, dz, batch, nin, nout, rows, scratch);
  mlp_wgrad_sum<<<blocks(np), 256, 0, st>>>(scratch, nchunks, np, dw);
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
