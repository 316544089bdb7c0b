/*
 * DBCSR stack multiply for gfx950: C[c_i] += A[a_i] * B[b_i] over a parameter stack.
 *
 * Semantics follow samples/smm/kernels/multiply.cl:12-15 with libxstream_common.h:86-87:
 * A block m-by-k column-major, B block k-by-n row-major (transposed by libsmm_acc_transpose),
 * C block m-by-n column-major; a stack record is ps ints whose first three are the a, b, c element
 * offsets with index base pz (multiply.cl:96-99: ps = 3, pz = 1 for param_format 0; ps = 6, pz = 0 for
 * DBM task records, seen from their fourth int).
 *
 * Work decomposition (new design, not the reference's mini-batch + atomic flush):
 *   - work-group w owns every C-run (maximal sequence of equal c) whose first entry lies in
 *     [w*bs, (w+1)*bs) and runs it to its end, so a C block is flushed once per run; a run longer
 *     than tsplit entries is cut at segment boundaries (lxs_cut) and its pieces flush atomically;
 *   - the threads cover one C block with TM x TN microtiles (ntiles threads); NT/ntiles such
 *     groups take the run's entries round-robin, each staging its A and B block in LDS;
 *   - partial C sums stay in registers for the whole run and are reduced across groups through
 *     LDS at the run boundary, then written with one read-modify-write (or one atomic add when
 *     the stack is not known to be grouped by C).
 *
 * This file is compiled once per data type (LXS_REAL/LXS_TAG) and registers its kernels in a table.
 */
#include <hip/hip_runtime.h>
#include "lxs_kernels.h"
#include "lxs_mfma.h"

#if !defined(LXS_REAL)
#  define LXS_REAL double
#  define LXS_TAG f64
#endif
#define LXS_CAT2(A, B) A##B
#define LXS_CAT(A, B) LXS_CAT2(A, B)

namespace {

typedef LXS_REAL real_t;

struct smm_entry_t {
  int tm, tn, nt, m, n, k; /* m = n = k = 0: runtime shape */
  int kt;                  /* compile-time k-slice (0: runtime, -1: wavefront-per-run kernel without LDS) */
  const void* fn;
};

template <typename T>
__device__ __forceinline__ void flush_one(T* p, T v, int atomic) {
  if (atomic) {
    unsafeAtomicAdd(p, v);
  }
  else {
    *p += v;
  }
}

template <typename T, int TM, int TN, int NT, int MC, int NC, int KC, int KTC>
__global__ void __launch_bounds__(NT) smm_stack(T* __restrict__ cdata, const T* __restrict__ adata,
  const T* __restrict__ bdata, const int* __restrict__ stack, int stack_size, int bs, int m_rt, int n_rt, int k_rt,
  int kt_rt, int atomic, int ps, int pz, int tsplit) {
  const int M = (0 < MC ? MC : m_rt), N = (0 < NC ? NC : n_rt), K = (0 < KC ? KC : k_rt);
  const int KT = (0 < KTC ? KTC : kt_rt);
  constexpr bool KEVEN = (0 < KC && 0 < KTC && 0 == (KC % KTC));
  const int tiles_m = (M + TM - 1) / TM, tiles_n = (N + TN - 1) / TN;
  const int ntiles = tiles_m * tiles_n;
  const int G = NT / ntiles; /* >= 1, checked on the host */
  const int MP = tiles_m * TM, NP = tiles_n * TN;
  const int tid = threadIdx.x;
  const int g = tid / ntiles, t = tid - g * ntiles;
  const bool active = (g < G);
  const int m0 = (t % tiles_m) * TM, n0 = (t / tiles_m) * TN;
  const int slab = KT * (MP + NP); /* LDS elements per group */
  extern __shared__ double lxs_smem[];
  T* const smem = reinterpret_cast<T*>(lxs_smem);
  T* const As = smem + (active ? g : 0) * slab; /* KT x MP, column-major (ld = MP) */
  T* const Bs = As + KT * MP;                   /* KT x NP, row-major (ld = NP) */
  __shared__ int s_idx;

  const int begin = blockIdx.x * bs;
  const int end = min(begin + bs, stack_size);

  /* padding rows/columns of the staged blocks are read by padded microtiles: keep them finite */
  for (int x = tid; x < G * slab; x += NT) smem[x] = T(0);
  if (0 == tid) s_idx = end;
  __syncthreads();
  /* first piece head in [begin, end): entry i starts a run if its c differs from entry i-1, and
     begin starts a piece where it cuts a long run */
  for (int i = begin + tid; i < end; i += NT) {
    const int ci = stack[ps * i + 2];
    if (0 == i || ci != stack[ps * (i - 1) + 2] || (i == begin && lxs::lxs_cut(stack, ps, i, ci, tsplit))) {
      atomicMin(&s_idx, i);
    }
  }
  __syncthreads();
  int h = s_idx;
  __syncthreads();
  if (h >= end) return; /* the whole range continues a run owned by an earlier work-group */

  for (;;) {
    const int c0 = stack[ps * h + 2];
    /* end of this piece: first entry after h with a different c, or a segment boundary past end
       that cuts the run (or stack_size) */
    if (0 == tid) s_idx = stack_size;
    __syncthreads();
    int r_end;
    for (int base = h + 1;; base += NT) {
      const int i = base + tid;
      if (i < stack_size && (stack[ps * i + 2] != c0 ||
                              (end <= i && 0 == (i - end) % bs && lxs::lxs_cut(stack, ps, i, c0, tsplit)))) {
        atomicMin(&s_idx, i);
      }
      __syncthreads();
      r_end = s_idx;
      __syncthreads();
      if (r_end < base + NT || base + NT >= stack_size) break;
    }

    /* summation order: each entry's k-sum starts from zero (prod); entries are summed in
       segments of SEG_LEN (seg) that are folded into the run total (acc). The longest rounding
       chain is then k + SEG_LEN + run/SEG_LEN instead of run*k, which keeps long C-runs accurate;
       microtiles above 16 elements skip the segment level to save registers. */
    constexpr bool SEGMENTED = (TM * TN <= 16);
    constexpr int SEG_LEN = 16;
    T acc[TM][TN], seg[TM][TN];
#pragma unroll
    for (int r = 0; r < TM; ++r) {
#pragma unroll
      for (int c = 0; c < TN; ++c) acc[r][c] = seg[r][c] = T(0);
    }
    int nsteps = 0;

    for (int sb = h; sb < r_end; sb += G) {
      const int e = sb + g;
      const bool valid = active && (e < r_end);
      const T* ablk = adata;
      const T* bblk = bdata;
      if (valid) {
        ablk = adata + (stack[ps * e + 0] - pz);
        bblk = bdata + (stack[ps * e + 1] - pz);
      }
      T prod[TM][TN];
#pragma unroll
      for (int r = 0; r < TM; ++r) {
#pragma unroll
        for (int c = 0; c < TN; ++c) prod[r][c] = T(0);
      }
      for (int kc = 0; kc < K; kc += KT) {
        const int kl = (KEVEN ? KTC : min(KT, K - kc));
        if (valid) {
          const T* const asrc = ablk + kc * M;
          const int na = kl * M;
          for (int x = t; x < na; x += ntiles) {
            const int kk = x / M, m = x - kk * M;
            As[kk * MP + m] = asrc[x];
          }
          const T* const bsrc = bblk + kc * N;
          const int nb = kl * N;
          for (int x = t; x < nb; x += ntiles) {
            const int kk = x / N, n = x - kk * N;
            Bs[kk * NP + n] = bsrc[x];
          }
        }
        __syncthreads();
        if (valid) {
#pragma unroll 4
          for (int kk = 0; kk < kl; ++kk) {
            T a[TM], b[TN];
#pragma unroll
            for (int r = 0; r < TM; ++r) a[r] = As[kk * MP + m0 + r];
#pragma unroll
            for (int c = 0; c < TN; ++c) b[c] = Bs[kk * NP + n0 + c];
#pragma unroll
            for (int r = 0; r < TM; ++r) {
#pragma unroll
              for (int c = 0; c < TN; ++c) prod[r][c] = fma(a[r], b[c], prod[r][c]);
            }
          }
        }
        __syncthreads();
      }
#pragma unroll
      for (int r = 0; r < TM; ++r) {
#pragma unroll
        for (int c = 0; c < TN; ++c) {
          if (SEGMENTED) {
            seg[r][c] += prod[r][c];
          }
          else {
            acc[r][c] += prod[r][c];
          }
        }
      }
      if (SEGMENTED && 0 == (++nsteps % SEG_LEN)) {
#pragma unroll
        for (int r = 0; r < TM; ++r) {
#pragma unroll
          for (int c = 0; c < TN; ++c) {
            acc[r][c] += seg[r][c];
            seg[r][c] = T(0);
          }
        }
      }
    }
    if (SEGMENTED) {
#pragma unroll
      for (int r = 0; r < TM; ++r) {
#pragma unroll
        for (int c = 0; c < TN; ++c) acc[r][c] += seg[r][c];
      }
    }

    T* const cblk = cdata + (c0 - pz);
    /* a piece of a cut run shares its C block with another work-group */
    const int fl = (atomic || (0 < h && stack[ps * (h - 1) + 2] == c0) ||
                    (r_end < stack_size && stack[ps * r_end + 2] == c0)) ? 1 : 0;
    if (1 == G) {
      if (active) {
#pragma unroll
        for (int c = 0; c < TN; ++c) {
#pragma unroll
          for (int r = 0; r < TM; ++r) {
            const int m = m0 + r, n = n0 + c;
            if (m < M && n < N) flush_one(cblk + n * M + m, acc[r][c], fl);
          }
        }
      }
    }
    else { /* reduce the groups' partial C blocks through LDS, in group order */
      const int per = ntiles * TM * TN;
      T* const red = smem;
      if (active) {
#pragma unroll
        for (int r = 0; r < TM; ++r) {
#pragma unroll
          for (int c = 0; c < TN; ++c) red[g * per + (t * TM + r) * TN + c] = acc[r][c];
        }
      }
      __syncthreads();
      for (int x = tid; x < per; x += NT) {
        T s = red[x];
        for (int gg = 1; gg < G; ++gg) s += red[gg * per + x];
        const int tt = x / (TM * TN), rc = x - tt * (TM * TN);
        const int r = rc / TN, c = rc - r * TN;
        const int m = (tt % tiles_m) * TM + r, n = (tt / tiles_m) * TN + c;
        if (m < M && n < N) flush_one(cblk + n * M + m, s, fl);
      }
      __syncthreads();
    }
    h = r_end;
    if (h >= end || h >= stack_size) break;
  }
}

/*
 * Tiny blocks (m n <= 64), compile-time shape: one wavefront per C-run and no LDS or barriers. Lane l is
 * entry slot l / (m n) and C element l % (m n) (column-major i, j); the P = 64 / (m n) slots take the
 * run's entries round-robin, each lane forming its element's dot products straight from the (cached) A
 * and B blocks, an entry's k-sum from zero and then added to the lane's run total. The slots are summed in
 * slot order by lane shuffles at the run end and the C block is updated once (or atomically). Wavefront w
 * of a work-group owns the runs that start in its part of the work-group's bs entries.
 */
template <typename T, int NT, int MC, int NC, int KC>
__global__ void __launch_bounds__(NT) smm_lanes(T* __restrict__ cdata, const T* __restrict__ adata,
  const T* __restrict__ bdata, const int* __restrict__ stack, int stack_size, int bs, int m_rt, int n_rt, int k_rt,
  int kt_rt, int atomic, int ps, int pz, int tsplit) {
  constexpr int MN = MC * NC, P = 64 / MN, W = NT / 64;
  static_assert(0 < MN && MN <= 64 && 0 == NT % 64, "one C block per wavefront slot");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int slot = lane / MN, el = lane - slot * MN;
  const bool live = (slot < P);
  const int i = el % MC, j = el / MC;
  const int bsw = (bs + W - 1) / W;
  const int wbeg = blockIdx.x * bs + wave * bsw;
  const int wend = min(min(wbeg + bsw, (int)(blockIdx.x + 1) * bs), stack_size);
  (void)m_rt, (void)n_rt, (void)k_rt, (void)kt_rt;
  if (wbeg >= wend) return;
  /* pieces: runs are cut at the wavefront segment starts q*bs + w*bsw (lxs_cut) */
  const auto boundary = [&](int x) { const int r = x % bs; return 0 == r % bsw; };
  int h = wbeg; /* the first piece that starts in [wbeg, wend), 64 entries per step */
  for (;;) {
    const int x = h + lane;
    bool start = false;
    if (x < wend) {
      const int cx = stack[ps * x + 2];
      start = (0 == x || cx != stack[ps * (x - 1) + 2] || (x == wbeg && lxs::lxs_cut(stack, ps, x, cx, tsplit)));
    }
    const unsigned long long bal = __ballot(start);
    if (0 != bal) {
      h += __ffsll((long long)bal) - 1;
      break;
    }
    h += 64;
    if (h >= wend) return;
  }
  while (h < wend) {
    const int c0 = stack[ps * h + 2];
    int r_end = h + 1; /* the piece's end, 64 entries per step */
    for (;;) {
      const int x = r_end + lane;
      bool stop = true;
      if (x < stack_size) {
        stop = (stack[ps * x + 2] != c0) || (wend <= x && boundary(x) && lxs::lxs_cut(stack, ps, x, c0, tsplit));
      }
      const unsigned long long bal = __ballot(stop);
      if (0 != bal) {
        r_end += __ffsll((long long)bal) - 1;
        break;
      }
      r_end += 64;
    }
    T acc = T(0);
    if (live) {
      /* U of the slot's entries per step, their loads issued together (entries past the run read the
         slot's first entry and are not added); the run total still adds the entries in order */
      constexpr int U = 4;
      for (int e = h + slot; e < r_end; e += U * P) {
        int ao[U], bo[U];
        T av[U][KC], bv[U][KC];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const int x = (e + u * P < r_end ? e + u * P : e);
          ao[u] = stack[ps * x] - pz;
          bo[u] = stack[ps * x + 1] - pz;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
#pragma unroll
          for (int kk = 0; kk < KC; ++kk) {
            av[u][kk] = adata[ao[u] + i + kk * MC];
            bv[u][kk] = bdata[bo[u] + kk * NC + j];
          }
        }
        __builtin_amdgcn_sched_barrier(0); /* every load of the step is issued before the first product */
#pragma unroll
        for (int u = 0; u < U; ++u) {
          T prod = T(0);
#pragma unroll
          for (int kk = 0; kk < KC; ++kk) prod = fma(av[u][kk], bv[u][kk], prod);
          if (e + u * P < r_end) acc += prod;
        }
      }
    }
    T tot = acc;
#pragma unroll
    for (int s = 1; s < P; ++s) tot += __shfl(acc, (lane + s * MN) & 63);
    const int fl = (atomic || (0 < h && stack[ps * (h - 1) + 2] == c0) ||
                    (r_end < stack_size && stack[ps * r_end + 2] == c0)) ? 1 : 0;
    if (0 == slot) flush_one(cdata + (c0 - pz) + i + j * MC, tot, fl);
    h = r_end;
  }
}

/* kt = -1 marks a wavefront-per-run kernel: launched without LDS */
#define LXS_LANES(M, N, K, TM, TN, NT) {TM, TN, NT, M, N, K, -1, (const void*)&smm_lanes<real_t, NT, M, N, K>},
#define LXS_GENERIC(TM, TN, NT) {TM, TN, NT, 0, 0, 0, 0, (const void*)&smm_stack<real_t, TM, TN, NT, 0, 0, 0, 0>},
#define LXS_SPECIAL(M, N, K, TM, TN, NT, KT) {TM, TN, NT, M, N, K, KT, (const void*)&smm_stack<real_t, TM, TN, NT, M, N, K, KT>},

const smm_entry_t smm_table[] = {
  /* runtime shape */
  LXS_GENERIC(1, 1, 64) LXS_GENERIC(1, 1, 128) LXS_GENERIC(1, 1, 256)
  LXS_GENERIC(2, 2, 64) LXS_GENERIC(2, 2, 128) LXS_GENERIC(2, 2, 256)
  LXS_GENERIC(3, 3, 64) LXS_GENERIC(3, 3, 128) LXS_GENERIC(3, 3, 256)
  LXS_GENERIC(4, 4, 64) LXS_GENERIC(4, 4, 128) LXS_GENERIC(4, 4, 256) LXS_GENERIC(4, 4, 512)
  LXS_GENERIC(2, 4, 128) LXS_GENERIC(4, 2, 128) LXS_GENERIC(5, 5, 256)
  /* compile-time shapes of the stack workloads */
  LXS_LANES(5, 5, 5, 1, 1, 64) LXS_LANES(5, 5, 5, 1, 1, 256)
  LXS_SPECIAL(13, 13, 13, 2, 2, 64, 13) LXS_SPECIAL(13, 13, 13, 2, 2, 256, 13) LXS_SPECIAL(13, 13, 13, 3, 3, 64, 13)
  LXS_SPECIAL(23, 23, 23, 3, 3, 64, 23) LXS_SPECIAL(23, 23, 23, 3, 3, 128, 23) LXS_SPECIAL(23, 23, 23, 3, 3, 256, 23)
  LXS_SPECIAL(23, 23, 23, 4, 4, 64, 23) LXS_SPECIAL(23, 23, 23, 2, 2, 256, 23)
  LXS_SPECIAL(32, 32, 32, 4, 4, 64, 32) LXS_SPECIAL(32, 32, 32, 4, 4, 128, 32) LXS_SPECIAL(32, 32, 32, 2, 2, 256, 32)
  LXS_SPECIAL(64, 64, 64, 4, 4, 256, 32) LXS_SPECIAL(64, 64, 64, 8, 8, 64, 32)
  LXS_SPECIAL(4, 9, 22, 1, 1, 64, 22) LXS_SPECIAL(4, 9, 22, 2, 3, 64, 22) LXS_SPECIAL(4, 9, 22, 1, 1, 256, 22)
  LXS_SPECIAL(45, 45, 13, 3, 3, 256, 13) LXS_SPECIAL(45, 45, 13, 4, 4, 256, 13) LXS_SPECIAL(45, 45, 13, 5, 5, 128, 13)
};

} /* namespace */

extern "C" const void* LXS_CAT(lxs_smm_table_, LXS_TAG)(int* count) {
  *count = (int)(sizeof(smm_table) / sizeof(*smm_table));
  return smm_table;
}
