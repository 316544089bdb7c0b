/*
 * Batched stack multiply: many stacks of different (m, n, k) in one persistent kernel launch.
 *
 * Each stack keeps the semantics of libsmm_acc_process (samples/smm/kernels/multiply.cl:12-15).
 * The host (batch.c) cuts every stack into chunks of run-aligned entries and packs them into tasks:
 *   - blocks covered by at most 4 output tiles of 16x16: one wavefront per chunk, a task holds up
 *     to 4 chunks (one per wavefront of the work-group, possibly of different stacks);
 *   - larger blocks (m, n <= 64: up to 16 tiles): the 4 wavefronts of the work-group split the tiles
 *     of one chunk, each walking the same entries independently.
 * The persistent grid strides over the task list; no atomics or barriers are needed for scheduling.
 * The per-chunk work is the matrix-core scheme of smm_mfma.hip with runtime shapes: fragments
 * loaded straight into registers and refilled as soon as they have been consumed, C-runs summed in
 * registers (segments of 16 entries) and flushed once with an atomic add (stacks of a batch may
 * share C pools); runs longer than the stack's tsplit are cut at chunk boundaries (lxs_cut).
 */
#include <hip/hip_runtime.h>
#include "lxs_kernels.h"
#include "lxs_mfma.h"

namespace {

using lxs::gload;
using lxs::mfma;

constexpr int KB = 4;      /* k-steps (of 4) per register stage */
constexpr int SEG_LEN = 16; /* entries per summation segment */

/* one wavefront runs one chunk of one stack over NQ output tiles q0, q0 + qstride, ... */
template <typename T, int NQ>
__device__ void run_chunk(const lxs_batch_stack_t& d, int chunk, int q0, int qstride, int lane) {
  typedef typename mfma<T>::acc_t acc_t;
  const T* __restrict__ const adata = (const T*)d.a;
  const T* __restrict__ const bdata = (const T*)d.b;
  T* __restrict__ const cdata = (T*)d.c;
  const int* __restrict__ const stack = d.stack;
  const int M = d.m, N = d.n, K = d.k, TJ = d.tj, stack_size = d.stack_size;
  const int KS = (K + 3) >> 2, NS = (KS + KB - 1) / KB;
  const int li = lane & 15, lk = lane >> 4;
  const int begin = chunk * d.bs;
  const int end = min(begin + d.bs, stack_size);
  const T* const zeros = lxs::zeros<T>();
  int h = end;
  for (int base = begin; base < end; base += 64) {
    const int i = base + lane;
    bool head = false;
    if (i < end) {
      head = (0 == i) || (stack[3 * i + 2] != stack[3 * i - 1]) ||
             (i == begin && lxs::lxs_cut(stack, 3, i, stack[3 * i + 2], d.tsplit));
    }
    const unsigned long long mask = __ballot(head);
    if (0 != mask) {
      h = base + __ffsll((long long)mask) - 1;
      break;
    }
  }
  if (h >= end) return;

  int tri[NQ], trj[NQ], aoff[NQ], boff[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int tile = q0 + q * qstride;
    tri[q] = tile / TJ;
    trj[q] = tile - tri[q] * TJ;
    aoff[q] = min(16 * tri[q] + li, M - 1);
    boff[q] = min(16 * trj[q] + li, N - 1);
  }
  acc_t acc[NQ], prod[NQ];
  T fa[KB][NQ], fb[KB][NQ];
  const acc_t zero = {T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = prod[q] = zero;

#define LXS_LOAD_KSTEP(KBI, A0, B0, KSTEP) \
  do { \
    const int k_ = 4 * (KSTEP) + lk; \
    const bool kv_ = (k_ < K); \
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) { \
      fa[KBI][q] = gload(kv_ ? (adata + ((A0) + k_ * M + aoff[q])) : zeros); \
      fb[KBI][q] = gload(kv_ ? (bdata + ((B0) + k_ * N + boff[q])) : zeros); \
    } \
  } while (0)

  int e = h;
  int a0 = stack[3 * e + 0] - 1, b0 = stack[3 * e + 1] - 1, c0 = stack[3 * e + 2];
  bool have1 = (e + 1 < stack_size), have2 = (e + 2 < stack_size);
  int a1 = (have1 ? stack[3 * e + 3] - 1 : a0), b1 = (have1 ? stack[3 * e + 4] - 1 : b0);
  int c1 = (have1 ? stack[3 * e + 5] : -1);
  int a2 = (have2 ? stack[3 * e + 6] - 1 : a0), b2 = (have2 ? stack[3 * e + 7] - 1 : b0);
  int c2 = (have2 ? stack[3 * e + 8] : -1);
  /* a long run is cut at the chunk boundaries past end (lxs_cut); nbd: the next one */
  int nbd = end;
  const auto cut_at = [&](int i, int ci) -> bool {
    if (i != nbd) return false;
    nbd += d.bs;
    if (i >= stack_size) return false;
    return lxs::lxs_cut(stack, 3, i, ci, d.tsplit);
  };
  bool cut1 = cut_at(e + 1, c1);
  bool cut2 = cut_at(e + 2, c2);
  int s = 0;
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    if (kb < KS) LXS_LOAD_KSTEP(kb, a0, b0, kb);
  }
  for (;;) { /* one C-run (or piece of one) per iteration */
    int nent = 0;
    bool run_ends, own_next;
    for (;;) { /* one unit (entry, k-stage) per iteration */
      const bool same_entry = (s + 1 < NS);
      run_ends = !same_entry && (!have1 || c1 != c0 || cut1);
      own_next = have1 && ((c1 == c0 && !cut1) || e + 1 < end);
      const int na = ((same_entry || !own_next) ? a0 : a1), nb = ((same_entry || !own_next) ? b0 : b1);
      const int ns = (same_entry ? s + 1 : 0);
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        if (s * KB + kb < KS) {
#pragma unroll
          for (int q = 0; q < NQ; ++q) prod[q] = mfma<T>::mma(fa[kb][q], fb[kb][q], prod[q]);
        }
        if (ns * KB + kb < KS) LXS_LOAD_KSTEP(kb, na, nb, ns * KB + kb);
      }
      if (same_entry) {
        ++s;
        continue;
      }
      if (SEG_LEN == ++nent && !run_ends) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          acc[q] += prod[q];
          prod[q] = zero;
        }
        nent = 0;
      }
      if (run_ends || !own_next) break;
      ++e;
      a0 = a1;
      b0 = b1;
      a1 = a2;
      b1 = b2;
      c1 = c2;
      have1 = have2;
      cut1 = cut2;
      s = 0;
      have2 = (e + 2 < stack_size);
      a2 = (have2 ? stack[3 * e + 6] - 1 : a0);
      b2 = (have2 ? stack[3 * e + 7] - 1 : b0);
      c2 = (have2 ? stack[3 * e + 8] : -1);
      cut2 = cut_at(e + 2, c2);
    }
    {
      T* const cblk = cdata + (c0 - 1);
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const int j = 16 * trj[q] + li;
        const acc_t sum = acc[q] + prod[q];
        prod[q] = acc[q] = zero;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = 16 * tri[q] + mfma<T>::row(lane, r);
          if (i < M && j < N) unsafeAtomicAdd(cblk + j * M + i, sum[r]);
        }
      }
    }
    if (!own_next) break;
    ++e;
    a0 = a1;
    b0 = b1;
    c0 = c1;
    a1 = a2;
    b1 = b2;
    c1 = c2;
    have1 = have2;
    cut1 = cut2;
    s = 0;
    have2 = (e + 2 < stack_size);
    a2 = (have2 ? stack[3 * e + 6] - 1 : a0);
    b2 = (have2 ? stack[3 * e + 7] - 1 : b0);
    c2 = (have2 ? stack[3 * e + 8] : -1);
    cut2 = cut_at(e + 2, c2);
  }
#undef LXS_LOAD_KSTEP
}

template <typename T>
__global__ void __launch_bounds__(256) smm_batch(const lxs_batch_stack_t* __restrict__ desc,
  const lxs_batch_task_t* __restrict__ tasks, int ntasks) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int t = blockIdx.x; t < ntasks; t += gridDim.x) {
    const int team = tasks[t].team;
    const int item = (0 != team ? 0 : wave);
    const int si = tasks[t].stack[item];
    if (0 > si) continue;
    {
      const lxs_batch_stack_t d = desc[si];
      const int ntiles = d.ti * d.tj;
      const int q0 = (0 != team ? wave : 0), qstride = (0 != team ? 4 : 1);
      const int nq = (ntiles - q0 + qstride - 1) / qstride; /* tiles of this wavefront (1..4) */
      const int chunk = tasks[t].chunk[item];
      switch (nq) {
        case 1: run_chunk<T, 1>(d, chunk, q0, qstride, lane); break;
        case 2: run_chunk<T, 2>(d, chunk, q0, qstride, lane); break;
        case 3: run_chunk<T, 3>(d, chunk, q0, qstride, lane); break;
        case 4: run_chunk<T, 4>(d, chunk, q0, qstride, lane); break;
        default: break; /* 0 tiles (team of a block with fewer than 4 tiles): nothing to do */
      }
    }
  }
}

} /* namespace */

extern "C" int lxs_batch_launch(int dtype, const lxs_batch_stack_t* desc, const lxs_batch_task_t* tasks, int ntasks,
  int nblocks, void* hip_stream) {
  if (0 >= ntasks) return 0;
  if (0 >= nblocks) return LXS_ELIMIT;
  if (3 == dtype) {
    smm_batch<double><<<nblocks, 256, 0, (hipStream_t)hip_stream>>>(desc, tasks, ntasks);
  }
  else if (1 == dtype) {
    smm_batch<float><<<nblocks, 256, 0, (hipStream_t)hip_stream>>>(desc, tasks, ntasks);
  }
  else {
    return LXS_ENOVARIANT;
  }
  return (int)hipGetLastError();
}
