/*
 * DBCSR stack multiply on gfx950 matrix cores: C[c_i] += A[a_i] * B[b_i] with MFMA 16x16x4
 * (v_mfma_f64_16x16x4f64 for FP64, v_mfma_f32_16x16x4f32 for FP32).
 *
 * Same semantics as smm_stack.hip (samples/smm/kernels/multiply.cl:12-15: A m-by-k column-major,
 * B k-by-n row-major, C m-by-n column-major; a stack record is ps ints whose first three are the a, b, c
 * element offsets with index base pz: ps = 3, pz = 1 for DBCSR stacks, ps = 6, pz = 0 for DBM task
 * records seen from their fourth int, multiply.cl:96-99); a different design:
 *   - a wavefront covers the C block with 16x16 output tiles (TI x TJ tiles, padded); a
 *     work-group of NWI wavefronts splits the tile rows, each wavefront walking the same entries
 *     independently (no LDS, no barriers);
 *   - MFMA operands are loaded straight from global memory into registers: an A fragment is 16
 *     consecutive rows of 4 columns of the column-major A block, a B fragment 16 consecutive
 *     columns of 4 rows of the row-major B block, so every load is a set of 128-byte rows;
 *   - loads roll: as soon as a k-step's fragments have been fed to the matrix core, the same
 *     registers are refilled with the next unit of work (next k-stage, or the next stack entry),
 *     so one entry's loads are in flight while the previous entry is multiplied;
 *   - stack indices are wave-uniform scalar loads, fetched one entry ahead;
 *   - C-run ownership as in smm_stack.hip: a work-group owns the runs whose first entry lies in
 *     [w*bs, (w+1)*bs), sums the run in registers (segments of 16 entries in the matrix-core
 *     accumulator, added to a run total) and flushes it once; a run longer than tsplit entries is
 *     cut into pieces at the segment boundaries (lxs_cut), so no work-group walks a long run alone;
 *   - fragment loads carry no branches and no selects on loaded values (padding lanes pick their
 *     address instead), so the compiler can count outstanding loads exactly.
 *
 * Fragment and output-tile lane layouts: lxs_mfma.h.
 */
#include <hip/hip_runtime.h>
#include <climits>
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
struct mfma_entry_t {
  int ti, tj, nwi, kb, m, n, k; /* m = n = k = 0: runtime shape with m <= 16*ti, n <= 16*tj (b4: 8*ti, 8*tj) */
  int ep;                       /* 1: smm_mfma; > 1: smm_mfma_ep with ep wavefronts per C-run */
  int b4;                       /* 1: smm_mfma_eb (4x4 blocks, ti x tj instructions per k-step), 2: smm_mfma_ek */
  const void* fn;
};

using lxs::gload;
using lxs::lxs_cut;
using lxs::mfma;

template <typename T, int TI, int TJ, int NWI, int KB, int MC, int NC, int KC>
__global__ void __launch_bounds__(64 * NWI) smm_mfma(T* __restrict__ cdata, const T* __restrict__ adata,
  const T* __restrict__ bdata, const int* __restrict__ stack, int stack_size, int bs, int m_rt, int n_rt, int k_rt,
  int kt_unused, int atomic, int ps, int pz, int tsplit) {
  typedef typename mfma<T>::acc_t acc_t;
  constexpr int RI = TI / NWI; /* tile rows per wavefront */
  /* summation order: the matrix core accumulates up to SEG_LEN entries in prod, which is then
     added to the run total acc; the longest rounding chain is SEG_LEN*k/4 + run/SEG_LEN */
  constexpr int SEG_LEN = 16;
  const int M = (0 < MC ? MC : m_rt), N = (0 < NC ? NC : n_rt), K = (0 < KC ? KC : k_rt);
  const int KS = (K + 3) >> 2;       /* k-steps of 4 */
  const int NS = (KS + KB - 1) / KB; /* k-stages per entry */
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 15, lk = lane >> 4;
  const int ti0 = wave * RI;
  (void)kt_unused;

  const int begin = blockIdx.x * bs;
  const int end = min(begin + bs, stack_size);
  /* first piece head in [begin, end): a run head, or begin where it cuts a long run */
  int h = end;
  for (int base = begin; base < end; base += 64) {
    const int i = base + lane;
    bool head = false;
    if (i < end) {
      head = (0 == i) || (stack[ps * i + 2] != stack[ps * (i - 1) + 2]) ||
             (i == begin && lxs_cut(stack, ps, i, stack[ps * i + 2], tsplit));
    }
    const unsigned long long mask = __ballot(head);
    if (0 != mask) {
      h = base + __ffsll((long long)mask) - 1;
      break;
    }
  }
  if (h >= end) return;

  /* per-lane parts of the fragment addresses. Lanes past m (A) or past n (B) only feed output
     rows/columns that are never stored, so they load a clamped, valid element of the block.
     Lanes past k must contribute zero to valid outputs: they load from a zeroed device array.
     The choice is made on the address, so no select waits on a loaded value. */
  int aoff[RI], boff[TJ];
#pragma unroll
  for (int ri = 0; ri < RI; ++ri) aoff[ri] = min(16 * (ti0 + ri) + li, M - 1);
#pragma unroll
  for (int tj = 0; tj < TJ; ++tj) boff[tj] = min(16 * tj + li, N - 1);
  const T* const zeros = lxs::zeros<T>();

  acc_t acc[RI][TJ], prod[RI][TJ];
  T fa[KB][RI], fb[KB][TJ];
  const acc_t zero = {T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int ri = 0; ri < RI; ++ri) {
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) acc[ri][tj] = prod[ri][tj] = zero;
  }

#define LXS_LOAD_KSTEP(KBI, A0, B0, KSTEP) \
  do { \
    const int k_ = 4 * (KSTEP) + lk; \
    const bool kv_ = (k_ < K); \
    _Pragma("unroll") for (int ri = 0; ri < RI; ++ri) { \
      fa[KBI][ri] = gload(kv_ ? (adata + ((A0) + k_ * M + aoff[ri])) : zeros); \
    } \
    _Pragma("unroll") for (int tj = 0; tj < TJ; ++tj) { \
      fb[KBI][tj] = gload(kv_ ? (bdata + ((B0) + k_ * N + boff[tj])) : zeros); \
    } \
  } while (0)

  /* stack indices: current entry e, next entry (1) and the one after (2), as scalar loads */
  int e = h;
  int a0 = stack[ps * e + 0] - pz, b0 = stack[ps * e + 1] - pz, c0 = stack[ps * e + 2];
  bool have1 = (e + 1 < stack_size), have2 = (e + 2 < stack_size);
  int a1 = (have1 ? stack[ps * (e + 1) + 0] - pz : a0), b1 = (have1 ? stack[ps * (e + 1) + 1] - pz : b0);
  int c1 = (have1 ? stack[ps * (e + 1) + 2] : -1);
  int a2 = (have2 ? stack[ps * (e + 2) + 0] - pz : a0), b2 = (have2 ? stack[ps * (e + 2) + 1] - pz : b0);
  int c2 = (have2 ? stack[ps * (e + 2) + 2] : -1);
  /* cuts past end: entry i >= end starts a new piece of a run when i is a segment boundary (next
     boundary: nbd) and lxs_cut holds; entries are checked once each, in order (cut1: e+1, cut2: e+2) */
  int nbd = end;
  const auto cut_at = [&](int i, int ci) -> bool {
    if (i != nbd) return false;
    nbd += bs;
    if (i >= stack_size) return false;
    return lxs_cut(stack, ps, i, ci, tsplit);
  };
  bool cut1 = cut_at(e + 1, c1);
  bool cut2 = cut_at(e + 2, c2);
  /* a piece that starts or ends inside a run shares its C block with another work-group */
  bool split = (0 < e && stack[ps * (e - 1) + 2] == c0);
  int s = 0;
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) {
    if (kb < KS) LXS_LOAD_KSTEP(kb, a0, b0, kb);
  }

  for (;;) { /* one C-run per iteration */
    int nent = 0;
    bool run_ends, own_next;
    for (;;) { /* one unit (entry, k-stage) per iteration; no memory traffic but the fragment loads */
      const bool same_entry = (s + 1 < NS);
      run_ends = !same_entry && (!have1 || c1 != c0 || cut1);
      own_next = have1 && ((c1 == c0 && !cut1) || e + 1 < end);
      /* the next unit of work; without one, reload the current entry (in bounds, never used) */
      const int na = ((same_entry || !own_next) ? a0 : a1), nb = ((same_entry || !own_next) ? b0 : b1);
      const int ns = (same_entry ? s + 1 : 0);
#pragma unroll
      for (int kb = 0; kb < KB; ++kb) {
        const int ks = s * KB + kb;
        if (ks < KS) {
#pragma unroll
          for (int ri = 0; ri < RI; ++ri) {
#pragma unroll
            for (int tj = 0; tj < TJ; ++tj) prod[ri][tj] = mfma<T>::mma(fa[kb][ri], fb[kb][tj], prod[ri][tj]);
          }
        }
        /* refill this slot with the next unit's fragments while the matrix core works */
        if (ns * KB + kb < KS) LXS_LOAD_KSTEP(kb, na, nb, ns * KB + kb);
      }
      if (same_entry) {
        ++s;
        continue;
      }
      if (SEG_LEN == ++nent && !run_ends) { /* fold the segment into the run total */
#pragma unroll
        for (int ri = 0; ri < RI; ++ri) {
#pragma unroll
          for (int tj = 0; tj < TJ; ++tj) {
            acc[ri][tj] += prod[ri][tj];
            prod[ri][tj] = zero;
          }
        }
        nent = 0;
      }
      if (run_ends || !own_next) break;
      ++e; /* next entry of the same run */
      a0 = a1;
      b0 = b1;
      a1 = a2;
      b1 = b2;
      c1 = c2;
      have1 = have2;
      cut1 = cut2;
      s = 0;
      have2 = (e + 2 < stack_size);
      a2 = (have2 ? stack[ps * (e + 2) + 0] - pz : a0);
      b2 = (have2 ? stack[ps * (e + 2) + 1] - pz : b0);
      c2 = (have2 ? stack[ps * (e + 2) + 2] : -1);
      cut2 = cut_at(e + 2, c2);
    }
    /* piece done: fold and flush C once (read-modify-write; atomic add for unsorted stacks and for
       pieces of a cut run) */
    {
      T* const cblk = cdata + (c0 - pz);
      const bool flush_atomic = atomic || split || (have1 && c1 == c0);
#pragma unroll
      for (int ri = 0; ri < RI; ++ri) {
#pragma unroll
        for (int tj = 0; tj < TJ; ++tj) {
          const int j = 16 * tj + li;
          const acc_t sum = acc[ri][tj] + prod[ri][tj];
          prod[ri][tj] = acc[ri][tj] = zero;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = 16 * (ti0 + ri) + mfma<T>::row(lane, r);
            if (i < M && j < N) {
              T* const p = cblk + j * M + i;
              if (flush_atomic) {
                unsafeAtomicAdd(p, sum[r]);
              }
              else {
                *p += sum[r];
              }
            }
          }
        }
      }
    }
    if (!own_next) break;
    ++e; /* first entry of the next run */
    a0 = a1;
    b0 = b1;
    c0 = c1;
    a1 = a2;
    b1 = b2;
    c1 = c2;
    have1 = have2;
    cut1 = cut2;
    split = false;
    s = 0;
    have2 = (e + 2 < stack_size);
    a2 = (have2 ? stack[ps * (e + 2) + 0] - pz : a0);
    b2 = (have2 ? stack[ps * (e + 2) + 1] - pz : b0);
    c2 = (have2 ? stack[ps * (e + 2) + 2] : -1);
    cut2 = cut_at(e + 2, c2);
  }
#undef LXS_LOAD_KSTEP
}

/*
 * Entry-parallel form for short runs: the EP wavefronts of a work-group share each C-run. Wavefront q
 * takes the run's entries h+q, h+q+EP, ... (each wavefront covers the whole C block with TI x TJ tiles, the
 * same rolling fragment loads as smm_mfma), then the EP partial blocks meet in LDS and are added in a
 * fixed order (wavefront 0 first), the tiles of the block shared out among the wavefronts for that sum
 * and the flush. So a run that is not cut gives the same bits on every call whatever the timing, and a
 * run of L entries costs each wavefront about L/EP entries instead of L (runs of ~30 entries and a few
 * thousand of them leave one wavefront per run too little memory parallelism).
 * Run ownership and cuts as in smm_mfma: the work-group owns the runs whose first entry lies in
 * [w*bs, (w+1)*bs); a run continues past that range until its C block changes or a segment boundary
 * past it cuts the run (lxs_cut). The end of a run is found by a 64-wide ballot over the C indices
 * (every wavefront computes it, so all take the same path through the barriers).
 * Barriers wait for LDS only (a fence restricted to the local address space), so fragment loads of the
 * next run stay in flight across them.
 *
 * B4 (FP64, smm_mfma_eb): the same run protocol on 4x4 blocks instead of 16x16 tiles. v_mfma_f64_4x4x4_4b
 * multiplies four independent 4x4x4 blocks at the FP64 matrix-core rate of the 16x16x4 form (a quarter of
 * its cycles), so a C block is padded to multiples of 8 instead of 16: 23 x 23 takes 24 x 24 (54 MFMA per
 * entry of 23 in k, 16 cycles each) instead of 32 x 32 (24 MFMA of 64 cycles), 1.78x less matrix-core time.
 * Lane l feeds k = l / 16 of block g = (l / 4) % 4, row l % 4 of its A block and column l % 4 of its B block;
 * the result of block g for A row i and B column j is in lane 16 i + 4 g + j (checked on gfx950 with a
 * probe). Block g covers the quadrant (g % 2, g / 2) of the padded 8 TI x 8 TJ C block, and instruction
 * (r, c) of a k-step takes its block rows r and columns c with the rows interleaved: block row i of
 * instruction r is C row 4 TI (g % 2) + TI i + r (columns likewise), so each lane's TI rows of A (TJ columns
 * of B) are consecutive in memory and load as pairs.
 * B4 = 2 (smm_mfma_ek, FP64 and FP32): the blocks of an instruction are one 4x4 block of C over consecutive k
 * of a 16-wide k-step: in FP64 four blocks of 4 k (block g takes k = 16 s + 4 g + l / 16), in FP32 the sixteen
 * blocks of v_mfma_f32_4x4x1_16b one k each (k = 16 s + l / 4; lxs::mfma4). A C block up to 4 TI x 4 TJ takes
 * TI x TJ instructions per 16 of k and every lane loads distinct elements (C row TI i + r of instruction r,
 * i = l % 4). The k-blocks are added in a fixed order where a segment is folded into the run total (FP64: lanes
 * l, l ^ 4, l ^ 8, l ^ 12, keeping element (i, j) in lane 16 i + j; FP32: lanes l ^ 4 .. l ^ 32, keeping row i
 * of column j in register i of lane j).
 */
template <typename T, int TI, int TJ, int EP, int KB, int MC, int NC, int KC, int B4>
__device__ __forceinline__ void ep_body(T* __restrict__ cdata, const T* __restrict__ adata,
  const T* __restrict__ bdata, const int* __restrict__ stack, int stack_size, int bs, int m_rt, int n_rt, int k_rt,
  int kt_unused, int atomic, int ps, int pz, int tsplit) {
  typedef typename mfma<T>::acc_t acc_t;
  constexpr int SEG_LEN = 16; /* entries per matrix-core accumulation before the fold into the run total */
  constexpr int NTILE = TI * TJ; /* 16x16 tiles, or (B4) 4x4-block instructions per k-step */
  const int M = (0 < MC ? MC : m_rt), N = (0 < NC ? NC : n_rt), K = (0 < KC ? KC : k_rt);
  const int KS = (2 == B4 ? (K + 15) >> 4 : (K + 3) >> 2); /* k-steps of 4 (B4 = 2: of 16, 4 per block) */
  const int NS = (KS + KB - 1) / KB; /* k-stages per entry */
  const int lane = threadIdx.x & 63;
  const int q = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int li = lane & 15, lk = lane >> 4;
  extern __shared__ double lxs_smem[];
  /* partial blocks: [EP][NTILE][4][64] (16x16 tiles), [EP][NTILE][64] (B4 = 1), [EP][NTILE][16] (B4 = 2) */
  T* const red = reinterpret_cast<T*>(lxs_smem);
  (void)kt_unused;

  const int begin = blockIdx.x * bs;
  const int end = min(begin + bs, stack_size);
  /* WIN: the stack around the work-group's entries is held in registers: lane l has the C, A and B offsets of
     entry begin + l and the C offset of entry begin - 64 + l (NONE outside the stack). The head and run-end
     scans, the cut tests and the entries' A and B offsets then come from these registers (shuffles and
     readlanes in wave-uniform code) instead of chains of dependent loads; entries outside the window are
     loaded. Kernels of 16 tiles or more keep loading: their accumulators leave no registers for the window
     (a spilled window register reloaded under a partial EXEC mask would feed the readlanes stale lanes). */
  constexpr bool WIN = (NTILE < 16) || (2 == B4 && NTILE <= 16); /* B4 = 2: one register pair per instruction */
  constexpr int NONE = INT_MIN; /* no C block: differs from every offset */
  const int wi = begin + lane, pi = begin - 64 + lane;
  const bool win = (wi < stack_size);
  int wc = NONE, wa = 0, wb = 0, pc = NONE;
  if constexpr (WIN) {
    wc = (win ? stack[ps * wi + 2] : NONE);
    wa = (win ? stack[ps * wi + 0] - pz : 0);
    wb = (win ? stack[ps * wi + 1] - pz : 0);
    pc = (0 <= pi ? stack[ps * pi + 2] : NONE);
  }
  /* C offset of a wave-uniform entry j (NONE outside the stack) */
  const auto c_uni = [&](int j) -> int {
    if constexpr (WIN) {
      if (begin <= j && j < begin + 64) return __builtin_amdgcn_readlane(wc, j - begin);
      if (begin - 64 <= j && j < begin) return __builtin_amdgcn_readlane(pc, j - begin + 64);
    }
    return (0 <= j && j < stack_size) ? stack[ps * j + 2] : NONE;
  };
  /* A and B offsets of a wave-uniform entry e of the stack */
  const auto a_uni = [&](int e) -> int {
    if constexpr (WIN) {
      if (begin <= e && e < begin + 64) return __builtin_amdgcn_readlane(wa, e - begin);
    }
    return stack[ps * e + 0] - pz;
  };
  const auto b_uni = [&](int e) -> int {
    if constexpr (WIN) {
      if (begin <= e && e < begin + 64) return __builtin_amdgcn_readlane(wb, e - begin);
    }
    return stack[ps * e + 1] - pz;
  };
  /* first piece head in [begin, end): a run head, or begin where it cuts a long run (lxs_cut) */
  int h = end;
  int base = begin;
  if constexpr (WIN) {
    /* C offset of the entry before each lane's: lanes 1..63 read lane l - 1 of wc, lane 0 reads lane 63 of
       the source, which holds entry begin - 1 (one shuffle, no lane-dependent select around a cross-lane read) */
    const int cprev = __shfl(63 == lane ? pc : wc, (lane + 63) & 63);
    const int ccut = ((0 < tsplit && tsplit <= begin) ? c_uni(begin - tsplit) : NONE);
    const bool head = (wi < end) && (0 == wi || wc != cprev || (wi == begin && ccut == wc));
    const unsigned long long mask = __ballot(head);
    if (0 != mask) h = begin + __ffsll((long long)mask) - 1;
    base = begin + 64;
  }
  for (; h >= end && base < end; base += 64) {
    const int i = base + lane;
    bool head = false;
    if (i < end) {
      head = (0 == i) || (stack[ps * i + 2] != stack[ps * (i - 1) + 2]) ||
             (i == begin && lxs_cut(stack, ps, i, stack[ps * i + 2], tsplit));
    }
    const unsigned long long mask = __ballot(head);
    if (0 != mask) h = base + __ffsll((long long)mask) - 1;
  }
  if (h >= end) return; /* the same for every wavefront of the work-group */
  /* end of the piece that starts at h: the first i > h with another C block, or a segment boundary past
     end where lxs_cut cuts the run */
  const auto run_end = [&](int h0, int c0) -> int {
    int from = h0 + 1;
    if constexpr (WIN) {
      if (h0 + 1 < begin + 64) {
        const int jc = wi - tsplit;
        const bool bnd = (end <= wi && 0 == (wi - end) % bs && 0 < tsplit && tsplit <= wi);
        const int x = __shfl(pc, (jc - begin + 64) & 63), y = __shfl(wc, (jc - begin) & 63);
        const bool need = (bnd && wc == c0); /* the cut test decides this lane */
        const bool inwin = (begin - 64 <= jc && jc < begin + 64);
        int cj = (jc < begin ? x : y); /* C offset of entry jc where the test needs it */
        if (need && !inwin) cj = stack[ps * jc + 2];
        const bool stop = (h0 < wi) && (!win || wc != c0 || (need && cj == c0));
        const unsigned long long mask = __ballot(stop);
        if (0 != mask) return begin + __ffsll((long long)mask) - 1;
        from = begin + 64;
      }
    }
    for (int b0 = from; b0 < stack_size; b0 += 64) {
      const int i = b0 + lane;
      bool stop = (i >= stack_size);
      if (!stop) {
        stop = (stack[ps * i + 2] != c0) || (end <= i && 0 == (i - end) % bs && lxs_cut(stack, ps, i, c0, tsplit));
      }
      const unsigned long long mask = __ballot(stop);
      if (0 != mask) return b0 + __ffsll((long long)mask) - 1;
    }
    return stack_size;
  };

  /* PAIR (even tile counts): paired tiles (lxs::pair_start), one two-element load per lane feeds two tiles,
     which halves the fragment load instructions and the cache lines they touch */
  constexpr bool PAIR = (0 == B4) && (0 == TI % 2) && (0 == TJ % 2);
  constexpr int PI = (PAIR ? TI / 2 : 1), PJ = (PAIR ? TJ / 2 : 1);
  typedef typename lxs::pair_of<T>::type v2_t;
  int aoff[TI], boff[TJ];
#pragma unroll
  for (int ti = 0; ti < TI; ++ti) aoff[ti] = (PAIR ? lxs::pair_start(ti >> 1, li, M) : min(16 * ti + li, M - 1));
#pragma unroll
  for (int tj = 0; tj < TJ; ++tj) boff[tj] = (PAIR ? lxs::pair_start(tj >> 1, li, N) : min(16 * tj + li, N - 1));
  const T* const zeros = lxs::zeros<T>();
  /* B4: this lane's first A row r4 and first B column c4 (TI rows, TJ columns from there); element pairs start
     at min(row, M - 2) so they stay inside the block, and a pair whose first row is the last row of the block
     (a4x, b4x) holds that row in its second element (the first is the row before; read at use, so that no
     select waits for a load in flight). Rows and columns past the block only feed padding outputs. */
  constexpr int TI2 = TI / 2, TJ2 = TJ / 2;
  const int g4 = (lane >> 2) & 3;
  const int r4 = (2 == B4 ? 0 : 4 * TI * (g4 & 1)) + TI * (lane & 3);
  const int c4 = (2 == B4 ? 0 : 4 * TJ * (g4 >> 1)) + TJ * (lane & 3);
  bool a4x[TI2 > 0 ? TI2 : 1], b4x[TJ2 > 0 ? TJ2 : 1];
#pragma unroll
  for (int p = 0; p < TI2; ++p) a4x[p] = (r4 + 2 * p == M - 1);
#pragma unroll
  for (int p = 0; p < TJ2; ++p) b4x[p] = (c4 + 2 * p == N - 1);
  v2_t qa[KB][TI2 > 0 ? TI2 : 1], qb[KB][TJ2 > 0 ? TJ2 : 1]; /* B4 element pairs */
  /* the matrix core accumulates up to SEG_LEN entries in prod; each segment is then added to the run total,
     which lives in this wavefront's partial block in LDS (red, slot q): total = 0 + prod first, total + prod
     after, the same additions a register total would make, without its registers */
  acc_t prod[TI][TJ];
  typedef typename lxs::mfma4<T>::acc_t acc4_t;
  acc4_t prod4[TI][TJ]; /* B4 */
  T fa[KB][TI], fb[KB][TJ];
  v2_t pa[KB][PI], pb[KB][PJ];
  const acc_t zero = {T(0), T(0), T(0), T(0)};
#pragma unroll
  for (int ti = 0; ti < TI; ++ti) {
#pragma unroll
    for (int tj = 0; tj < TJ; ++tj) {
      prod[ti][tj] = zero;
      prod4[ti][tj] = acc4_t{};
    }
  }
  bool folded = false; /* the run total in LDS holds a segment */
  const auto fold = [&]() {
#pragma unroll
    for (int ti = 0; ti < TI; ++ti) {
#pragma unroll
      for (int tj = 0; tj < TJ; ++tj) {
        if constexpr (2 == B4) {
          if constexpr (lxs::mfma4<T>::FP32) {
            /* sixteen k-blocks: lanes of one column j = l % 4 hold row i in register i, added over l ^ 4 .. l ^ 32 in
               a fixed order; lanes 0..3 keep element 4 i + j */
#pragma unroll
            for (int x = 0; x < 4; ++x) {
              T v = prod4[ti][tj][x];
              v += __shfl_xor(v, 4);
              v += __shfl_xor(v, 8);
              v += __shfl_xor(v, 16);
              v += __shfl_xor(v, 32);
              if (lane < 4) {
                T* const slot = red + (q * NTILE + ti * TJ + tj) * 16 + 4 * x + lane;
                *slot = (folded ? *slot : T(0)) + v;
              }
            }
          }
          else { /* the four k-quarters of the block, added in a fixed order, then lanes 16 i + j */
            const T v01 = prod4[ti][tj] + __shfl_xor(prod4[ti][tj], 4);
            const T v = v01 + __shfl_xor(v01, 8);
            if (0 == g4) {
              T* const slot = red + (q * NTILE + ti * TJ + tj) * 16 + ((lane >> 4) << 2) + (lane & 3);
              *slot = (folded ? *slot : T(0)) + v;
            }
          }
          prod4[ti][tj] = acc4_t{};
        }
        else if constexpr (1 == B4) {
          T* const slot = red + (q * NTILE + ti * TJ + tj) * 64 + lane;
          *slot = (folded ? *slot : T(0)) + prod4[ti][tj];
          prod4[ti][tj] = acc4_t{};
        }
        else {
#pragma unroll
          for (int x = 0; x < 4; ++x) {
            T* const slot = red + ((q * NTILE + ti * TJ + tj) * 4 + x) * 64 + lane;
            *slot = (folded ? *slot : T(0)) + prod[ti][tj][x];
          }
          prod[ti][tj] = zero;
        }
      }
    }
    folded = true;
  };

#define LXS_LOAD_KSTEP(KBI, A0, B0, KSTEP) \
  do { \
    const int k_ = (2 == B4 ? 16 * (KSTEP) + lxs::mfma4<T>::klane(lane) : 4 * (KSTEP) + lk); \
    const bool kv_ = (k_ < K); \
    if constexpr (0 != B4) { \
      _Pragma("unroll") for (int p_ = 0; p_ < TI2; ++p_) { \
        qa[KBI][p_] = lxs::gload2(kv_ ? (adata + ((A0) + k_ * M + min(r4 + 2 * p_, M - 2))) : zeros); \
      } \
      if constexpr (0 != TI % 2) fa[KBI][0] = gload(kv_ ? (adata + ((A0) + k_ * M + min(r4 + TI - 1, M - 1))) : zeros); \
      _Pragma("unroll") for (int p_ = 0; p_ < TJ2; ++p_) { \
        qb[KBI][p_] = lxs::gload2(kv_ ? (bdata + ((B0) + k_ * N + min(c4 + 2 * p_, N - 2))) : zeros); \
      } \
      if constexpr (0 != TJ % 2) fb[KBI][0] = gload(kv_ ? (bdata + ((B0) + k_ * N + min(c4 + TJ - 1, N - 1))) : zeros); \
    } \
    else if constexpr (PAIR) { \
      _Pragma("unroll") for (int p_ = 0; p_ < PI; ++p_) { \
        pa[KBI][p_] = lxs::gload2(kv_ ? (adata + ((A0) + k_ * M + aoff[2 * p_])) : zeros); \
      } \
      _Pragma("unroll") for (int p_ = 0; p_ < PJ; ++p_) { \
        pb[KBI][p_] = lxs::gload2(kv_ ? (bdata + ((B0) + k_ * N + boff[2 * p_])) : zeros); \
      } \
    } \
    else { \
      _Pragma("unroll") for (int ti = 0; ti < TI; ++ti) { \
        fa[KBI][ti] = gload(kv_ ? (adata + ((A0) + k_ * M + aoff[ti])) : zeros); \
      } \
      _Pragma("unroll") for (int tj = 0; tj < TJ; ++tj) { \
        fb[KBI][tj] = gload(kv_ ? (bdata + ((B0) + k_ * N + boff[tj])) : zeros); \
      } \
    } \
  } while (0)

  int e = -1, a0 = 0, b0 = 0; /* entry whose first k-stage is in flight in the fragment registers */
  for (;;) {                  /* one run (or piece of a cut run) [h, r) per iteration */
    const int c0 = c_uni(h);
    const int r = run_end(h, c0);
    /* a piece that starts or ends inside a run shares its C block with another work-group */
    const bool flush_atomic = atomic || (0 < h && c_uni(h - 1) == c0) || (r < stack_size && c_uni(r) == c0);
    const bool own_next = (r < end);
    int cur = h + q;
    if (cur < r) {
      if (cur != e) { /* not prefetched: load the first k-stage now */
        a0 = a_uni(cur);
        b0 = b_uni(cur);
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
          if (kb < KS) LXS_LOAD_KSTEP(kb, a0, b0, kb);
        }
      }
      int nent = 0;
      for (;;) { /* one entry of this wavefront per iteration */
        const int nx = cur + EP;
        const bool in_run = (nx < r);
        /* after the run: this wavefront's first entry of the next run, if the work-group owns it */
        const int cand = r + q;
        const bool pre = !in_run && own_next && cand < stack_size;
        const int nxt = (in_run ? nx : (pre ? cand : -1));
        const int na = (0 <= nxt ? a_uni(nxt) : a0), nb = (0 <= nxt ? b_uni(nxt) : b0);
        for (int s = 0; s < NS; ++s) {
          const bool last = (s + 1 == NS);
          const int la = (last ? na : a0), lb = (last ? nb : b0), ns = (last ? 0 : s + 1);
#pragma unroll
          for (int kb = 0; kb < KB; ++kb) {
            const int ks = s * KB + kb;
            if constexpr (0 != B4) {
              if (ks < KS) {
                T av[TI], bv[TJ];
#pragma unroll
                for (int p = 0; p < TI2; ++p) {
                  av[2 * p] = (a4x[p] ? qa[kb][p][1] : qa[kb][p][0]);
                  av[2 * p + 1] = qa[kb][p][1];
                }
                if constexpr (0 != TI % 2) av[TI - 1] = fa[kb][0];
#pragma unroll
                for (int p = 0; p < TJ2; ++p) {
                  bv[2 * p] = (b4x[p] ? qb[kb][p][1] : qb[kb][p][0]);
                  bv[2 * p + 1] = qb[kb][p][1];
                }
                if constexpr (0 != TJ % 2) bv[TJ - 1] = fb[kb][0];
#pragma unroll
                for (int ti = 0; ti < TI; ++ti) {
#pragma unroll
                  for (int tj = 0; tj < TJ; ++tj) {
                    prod4[ti][tj] = lxs::mfma4<T>::mma(av[ti], bv[tj], prod4[ti][tj]);
                  }
                }
              }
            }
            else if (ks < KS) {
#pragma unroll
              for (int ti = 0; ti < TI; ++ti) {
#pragma unroll
                for (int tj = 0; tj < TJ; ++tj) {
                  if constexpr (PAIR) {
                    prod[ti][tj] = mfma<T>::mma(pa[kb][ti >> 1][ti & 1], pb[kb][tj >> 1][tj & 1], prod[ti][tj]);
                  }
                  else {
                    prod[ti][tj] = mfma<T>::mma(fa[kb][ti], fb[kb][tj], prod[ti][tj]);
                  }
                }
              }
            }
            /* refill the slot with the next unit while the matrix core works */
            if (ns * KB + kb < KS) LXS_LOAD_KSTEP(kb, la, lb, ns * KB + kb);
          }
        }
        if (SEG_LEN == ++nent) { /* fold the segment into the run total */
          fold();
          nent = 0;
        }
        a0 = na;
        b0 = nb;
        if (!in_run) {
          e = nxt;
          break;
        }
        cur = nx;
      }
    }
    /* this wavefront's partial block (run total plus the open segment) in LDS, then the fixed-order sum and
       the flush, tiles shared out */
    fold();
    folded = false;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    if constexpr (2 == B4) { /* four instructions t at a time, one per quarter-wavefront: row i, column j */
      T* const cblk = cdata + (c0 - pz);
      const int x = lane & 15, rb = TI * (x >> 2), cb = TJ * (x & 3);
      for (int t0 = 4 * q; t0 < NTILE; t0 += 4 * EP) {
        const int t = t0 + (lane >> 4);
        if (t < NTILE) {
          const int r = t / TJ, c = t - r * TJ;
          T v = red[t * 16 + x];
#pragma unroll
          for (int p = 1; p < EP; ++p) v += red[(p * NTILE + t) * 16 + x];
          const int i = rb + r, j = cb + c;
          if (i < M && j < N) {
            T* const ptr = cblk + j * M + i;
            if (flush_atomic) {
              unsafeAtomicAdd(ptr, v);
            }
            else {
              *ptr += v;
            }
          }
        }
      }
    }
    else if constexpr (1 == B4) { /* instruction t = (r, c): this lane's block row lane / 16, column lane % 4 */
      T* const cblk = cdata + (c0 - pz);
      const int rb = 4 * TI * (g4 & 1) + TI * (lane >> 4), cb = 4 * TJ * (g4 >> 1) + TJ * (lane & 3);
      for (int t = q; t < NTILE; t += EP) {
        const int r = t / TJ, c = t - r * TJ;
        T v = red[t * 64 + lane];
#pragma unroll
        for (int p = 1; p < EP; ++p) v += red[(p * NTILE + t) * 64 + lane];
        const int i = rb + r, j = cb + c;
        if (i < M && j < N) {
          T* const ptr = cblk + j * M + i;
          if (flush_atomic) {
            unsafeAtomicAdd(ptr, v);
          }
          else {
            *ptr += v;
          }
        }
      }
    }
    else {
      T* const cblk = cdata + (c0 - pz);
      for (int t = q; t < NTILE; t += EP) {
        const int ti = t / TJ, tj = t - ti * TJ;
        /* element of the block behind output row/column l of a tile, and whether this lane writes it */
        const int j = (PAIR ? lxs::pair_start(tj >> 1, li, N) + (tj & 1) : 16 * tj + li);
        const bool jown = (PAIR ? lxs::pair_owner(tj >> 1, tj & 1, li, N) : j < N);
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          T v = red[(t * 4 + x) * 64 + lane];
#pragma unroll
          for (int p = 1; p < EP; ++p) v += red[((p * NTILE + t) * 4 + x) * 64 + lane];
          const int ii = mfma<T>::row(lane, x);
          const int i = (PAIR ? lxs::pair_start(ti >> 1, ii, M) + (ti & 1) : 16 * ti + ii);
          const bool iown = (PAIR ? lxs::pair_owner(ti >> 1, ti & 1, ii, M) : i < M);
          if (iown && jown) {
            T* const ptr = cblk + j * M + i;
            if (flush_atomic) {
              unsafeAtomicAdd(ptr, v);
            }
            else {
              *ptr += v;
            }
          }
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier(); /* every sum is read before the next run's partials overwrite LDS */
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    if (!own_next) break;
    h = r;
  }
#undef LXS_LOAD_KSTEP
}

#define LXS_EP_ARGS \
  T* __restrict__ cdata, const T* __restrict__ adata, const T* __restrict__ bdata, const int* __restrict__ stack, \
    int stack_size, int bs, int m_rt, int n_rt, int k_rt, int kt_unused, int atomic, int ps, int pz, int tsplit
#define LXS_EP_PASS cdata, adata, bdata, stack, stack_size, bs, m_rt, n_rt, k_rt, kt_unused, atomic, ps, pz, tsplit

template <typename T, int TI, int TJ, int EP, int KB, int MC, int NC, int KC>
__global__ void __launch_bounds__(64 * EP, (TI * TJ <= 4 ? 4 : 1)) smm_mfma_ep(LXS_EP_ARGS) {
  ep_body<T, TI, TJ, EP, KB, MC, NC, KC, 0>(LXS_EP_PASS);
}
/* FP64 on 4x4 blocks: C blocks up to 8 TI x 8 TJ */
template <typename T, int TI, int TJ, int EP, int KB, int MC, int NC, int KC>
__global__ void __launch_bounds__(64 * EP, 4) smm_mfma_eb(LXS_EP_ARGS) {
  ep_body<T, TI, TJ, EP, KB, MC, NC, KC, 1>(LXS_EP_PASS);
}
/* FP64 on 4x4 blocks that split k: C blocks up to 4 TI x 4 TJ, KB k-steps of 16 per stage */
template <typename T, int TI, int TJ, int EP, int KB, int MC, int NC, int KC>
__global__ void __launch_bounds__(64 * EP, 4) smm_mfma_ek(LXS_EP_ARGS) {
  ep_body<T, TI, TJ, EP, KB, MC, NC, KC, 2>(LXS_EP_PASS);
}

#define LXS_MG(TI, TJ, NWI, KB) {TI, TJ, NWI, KB, 0, 0, 0, 1, 0, (const void*)&smm_mfma<real_t, TI, TJ, NWI, KB, 0, 0, 0>},
#define LXS_MS(M, N, K, TI, TJ, NWI, KB) {TI, TJ, NWI, KB, M, N, K, 1, 0, (const void*)&smm_mfma<real_t, TI, TJ, NWI, KB, M, N, K>},
#define LXS_ME(M, N, K, TI, TJ, EP, KB) {TI, TJ, 1, KB, M, N, K, EP, 0, (const void*)&smm_mfma_ep<real_t, TI, TJ, EP, KB, M, N, K>},
#define LXS_MEG(TI, TJ, EP, KB) {TI, TJ, 1, KB, 0, 0, 0, EP, 0, (const void*)&smm_mfma_ep<real_t, TI, TJ, EP, KB, 0, 0, 0>},
#define LXS_EB(M, N, K, TI, TJ, EP, KB) {TI, TJ, 1, KB, M, N, K, EP, 1, (const void*)&smm_mfma_eb<real_t, TI, TJ, EP, KB, M, N, K>},
#define LXS_EBG(TI, TJ, EP, KB) {TI, TJ, 1, KB, 0, 0, 0, EP, 1, (const void*)&smm_mfma_eb<real_t, TI, TJ, EP, KB, 0, 0, 0>},
/* kb of a k-split kernel: 4 KB (its k per stage is 16 KB) */
#define LXS_EK(M, N, K, TI, TJ, EP, KB) {TI, TJ, 1, 4 * KB, M, N, K, EP, 2, (const void*)&smm_mfma_ek<real_t, TI, TJ, EP, KB, M, N, K>},
#define LXS_EKG(TI, TJ, EP, KB) {TI, TJ, 1, 4 * KB, 0, 0, 0, EP, 2, (const void*)&smm_mfma_ek<real_t, TI, TJ, EP, KB, 0, 0, 0>},
#define LXS_IS_f64 1 /* the 4x4-block kernels are FP64 only */

const mfma_entry_t mfma_table[] = {
  /* runtime shape (m <= 16*ti, n <= 16*tj), 4 k-steps (16 of k) per stage: every tile grid up to
     5x5 (m, n <= 80) with one wavefront per tile row, plus one wavefront for grids up to 4 tiles */
  LXS_MG(1, 1, 1, 4) LXS_MG(1, 2, 1, 4) LXS_MG(1, 3, 1, 4) LXS_MG(1, 4, 1, 4) LXS_MG(1, 5, 1, 4)
  LXS_MG(2, 1, 1, 4) LXS_MG(2, 1, 2, 4) LXS_MG(2, 2, 1, 4) LXS_MG(2, 2, 2, 4) LXS_MG(2, 3, 2, 4)
  LXS_MG(2, 3, 1, 2) LXS_MG(2, 4, 2, 4) LXS_MG(2, 5, 2, 4) LXS_MG(3, 1, 1, 4) LXS_MG(3, 1, 3, 4)
  LXS_MG(3, 2, 3, 4) LXS_MG(3, 2, 1, 2) LXS_MG(3, 3, 3, 4) LXS_MG(3, 4, 3, 4) LXS_MG(3, 5, 3, 4)
  LXS_MG(4, 1, 1, 4) LXS_MG(4, 1, 4, 4) LXS_MG(4, 2, 4, 4) LXS_MG(4, 3, 4, 4) LXS_MG(4, 4, 4, 4)
  LXS_MG(4, 5, 4, 4) LXS_MG(5, 1, 5, 4) LXS_MG(5, 1, 1, 2) LXS_MG(5, 2, 5, 4) LXS_MG(5, 3, 5, 4)
  LXS_MG(5, 4, 5, 4) LXS_MG(5, 5, 5, 4)
  /* compile-time shapes of the stack workloads (whole k per stage where registers allow) */
  LXS_MS(5, 5, 5, 1, 1, 1, 2) LXS_MS(13, 13, 13, 1, 1, 1, 4)
  LXS_MS(23, 23, 23, 2, 2, 1, 6) LXS_MS(23, 23, 23, 2, 2, 2, 6) LXS_MS(23, 23, 23, 2, 2, 1, 3)
  LXS_MS(32, 32, 32, 2, 2, 1, 8) LXS_MS(32, 32, 32, 2, 2, 2, 8) LXS_MS(32, 32, 32, 2, 2, 1, 4)
  LXS_MS(64, 64, 64, 4, 4, 4, 4) LXS_MS(64, 64, 64, 4, 4, 2, 4) LXS_MS(64, 64, 64, 4, 4, 4, 8)
  LXS_MS(4, 9, 22, 1, 1, 1, 6) LXS_MS(45, 45, 13, 3, 3, 1, 4) LXS_MS(45, 45, 13, 3, 3, 3, 4)
  /* entry-parallel (deterministic whole runs) for the stack workloads */
  LXS_ME(23, 23, 23, 2, 2, 2, 6) LXS_ME(23, 23, 23, 2, 2, 4, 6) LXS_ME(23, 23, 23, 2, 2, 8, 6)
  LXS_ME(13, 13, 13, 1, 1, 4, 4) LXS_ME(13, 13, 13, 1, 1, 8, 4) LXS_ME(32, 32, 32, 2, 2, 2, 8)
  LXS_ME(32, 32, 32, 2, 2, 4, 8) LXS_ME(64, 64, 64, 4, 4, 2, 4) LXS_ME(45, 45, 13, 3, 3, 2, 4)
  LXS_ME(5, 5, 5, 1, 1, 4, 2) LXS_ME(4, 9, 22, 1, 1, 4, 6)
  /* entry-parallel, runtime shape (m <= 16*ti, n <= 16*tj), 4 k-steps (16 of k) per stage: 2 wavefronts per run
     up to 16 tiles (the FP64 partial blocks fill 64 KB of LDS there), 4 up to 8 tiles, 8 up to 4 tiles */
  LXS_MEG(1, 1, 2, 4) LXS_MEG(1, 2, 2, 4) LXS_MEG(1, 3, 2, 4) LXS_MEG(1, 4, 2, 4) LXS_MEG(1, 5, 2, 4)
  LXS_MEG(2, 1, 2, 4) LXS_MEG(2, 2, 2, 4) LXS_MEG(2, 3, 2, 4) LXS_MEG(2, 4, 2, 4) LXS_MEG(2, 5, 2, 4)
  LXS_MEG(3, 1, 2, 4) LXS_MEG(3, 2, 2, 4) LXS_MEG(3, 3, 2, 4) LXS_MEG(3, 4, 2, 4) LXS_MEG(3, 5, 2, 4)
  LXS_MEG(4, 1, 2, 4) LXS_MEG(4, 2, 2, 4) LXS_MEG(4, 3, 2, 4) LXS_MEG(4, 4, 2, 4) LXS_MEG(5, 1, 2, 4)
  LXS_MEG(5, 2, 2, 4) LXS_MEG(5, 3, 2, 4) LXS_MEG(1, 1, 4, 4) LXS_MEG(1, 2, 4, 4) LXS_MEG(1, 3, 4, 4)
  LXS_MEG(1, 4, 4, 4) LXS_MEG(1, 5, 4, 4) LXS_MEG(2, 1, 4, 4) LXS_MEG(2, 2, 4, 4) LXS_MEG(2, 3, 4, 4)
  LXS_MEG(2, 4, 4, 4) LXS_MEG(3, 1, 4, 4) LXS_MEG(3, 2, 4, 4) LXS_MEG(4, 1, 4, 4) LXS_MEG(4, 2, 4, 4)
  LXS_MEG(5, 1, 4, 4) LXS_MEG(1, 1, 8, 4) LXS_MEG(1, 2, 8, 4) LXS_MEG(1, 3, 8, 4) LXS_MEG(1, 4, 8, 4)
  LXS_MEG(2, 1, 8, 4) LXS_MEG(2, 2, 8, 4) LXS_MEG(3, 1, 8, 4) LXS_MEG(4, 1, 8, 4)
#if LXS_CAT(LXS_IS_, LXS_TAG)
  /* entry-parallel on 4x4 blocks (FP64): the stack workloads whose blocks pad less to multiples of 8 than to 16 */
  LXS_EB(23, 23, 23, 3, 3, 4, 3) LXS_EB(23, 23, 23, 3, 3, 8, 3) LXS_EB(23, 23, 23, 3, 3, 4, 6)
  LXS_EB(23, 23, 23, 3, 3, 8, 6) LXS_EB(5, 5, 5, 1, 1, 4, 2) LXS_EB(5, 5, 5, 1, 1, 8, 2)
  LXS_EB(13, 13, 13, 2, 2, 4, 4) LXS_EB(13, 13, 13, 2, 2, 8, 4) LXS_EB(4, 9, 22, 1, 2, 4, 6)
  LXS_EB(4, 9, 22, 1, 2, 8, 6)
  /* runtime shape (m <= 8*ti, n <= 8*tj), 4 k-steps per stage (2 for 16 blocks), 4 or 8 wavefronts per run */
  LXS_EBG(1, 1, 4, 4) LXS_EBG(1, 2, 4, 4) LXS_EBG(1, 3, 4, 4) LXS_EBG(1, 4, 4, 4) LXS_EBG(2, 1, 4, 4)
  LXS_EBG(2, 2, 4, 4) LXS_EBG(2, 3, 4, 4) LXS_EBG(2, 4, 4, 4) LXS_EBG(3, 1, 4, 4) LXS_EBG(3, 2, 4, 4)
  LXS_EBG(3, 3, 4, 4) LXS_EBG(3, 4, 4, 4) LXS_EBG(4, 1, 4, 4) LXS_EBG(4, 2, 4, 4) LXS_EBG(4, 3, 4, 4)
  LXS_EBG(4, 4, 4, 2) LXS_EBG(1, 1, 8, 4) LXS_EBG(1, 2, 8, 4) LXS_EBG(1, 3, 8, 4) LXS_EBG(1, 4, 8, 4)
  LXS_EBG(2, 1, 8, 4) LXS_EBG(2, 2, 8, 4) LXS_EBG(2, 3, 8, 4) LXS_EBG(2, 4, 8, 4) LXS_EBG(3, 1, 8, 4)
  LXS_EBG(3, 2, 8, 4) LXS_EBG(3, 3, 8, 4) LXS_EBG(3, 4, 8, 4) LXS_EBG(4, 1, 8, 4) LXS_EBG(4, 2, 8, 4)
  LXS_EBG(4, 3, 8, 4) LXS_EBG(4, 4, 8, 2)
#endif
  /* entry-parallel on 4x4 blocks that split k (four quarters of 16 in FP64, sixteen single k in FP32): small
     blocks, one 16-k step (or two) per stage */
  LXS_EK(5, 5, 5, 2, 2, 4, 1) LXS_EK(5, 5, 5, 2, 2, 8, 1) LXS_EK(13, 13, 13, 4, 4, 4, 1) LXS_EK(13, 13, 13, 4, 4, 8, 1)
  LXS_EK(4, 9, 22, 1, 3, 4, 2) LXS_EK(4, 9, 22, 1, 3, 8, 2)
  /* runtime shape (m <= 4*ti, n <= 4*tj, up to 16 x 16) */
  LXS_EKG(1, 1, 4, 1) LXS_EKG(1, 2, 4, 1) LXS_EKG(1, 3, 4, 1) LXS_EKG(1, 4, 4, 1) LXS_EKG(2, 1, 4, 1) LXS_EKG(2, 2, 4, 1)
  LXS_EKG(2, 3, 4, 1) LXS_EKG(2, 4, 4, 1) LXS_EKG(3, 1, 4, 1) LXS_EKG(3, 2, 4, 1) LXS_EKG(3, 3, 4, 1) LXS_EKG(3, 4, 4, 1)
  LXS_EKG(4, 1, 4, 1) LXS_EKG(4, 2, 4, 1) LXS_EKG(4, 3, 4, 1) LXS_EKG(4, 4, 4, 1) LXS_EKG(1, 1, 8, 1) LXS_EKG(1, 2, 8, 1)
  LXS_EKG(1, 3, 8, 1) LXS_EKG(1, 4, 8, 1) LXS_EKG(2, 1, 8, 1) LXS_EKG(2, 2, 8, 1) LXS_EKG(2, 3, 8, 1) LXS_EKG(2, 4, 8, 1)
  LXS_EKG(3, 1, 8, 1) LXS_EKG(3, 2, 8, 1) LXS_EKG(3, 3, 8, 1) LXS_EKG(3, 4, 8, 1) LXS_EKG(4, 1, 8, 1) LXS_EKG(4, 2, 8, 1)
  LXS_EKG(4, 3, 8, 1) LXS_EKG(4, 4, 8, 1) LXS_EKG(1, 1, 4, 2) LXS_EKG(1, 2, 4, 2) LXS_EKG(1, 3, 4, 2) LXS_EKG(1, 4, 4, 2)
  LXS_EKG(2, 1, 4, 2) LXS_EKG(2, 2, 4, 2) LXS_EKG(3, 1, 4, 2) LXS_EKG(4, 1, 4, 2)
};

} /* namespace */

extern "C" const void* LXS_CAT(lxs_mfma_table_, LXS_TAG)(int* count) {
  *count = (int)(sizeof(mfma_table) / sizeof(*mfma_table));
  return mfma_table;
}
