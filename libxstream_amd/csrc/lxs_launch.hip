/*
 * Launchers for the gfx950 kernels: stack-kernel dispatch over the per-type kernel tables,
 * in-place block transpose (samples/smm/kernels/transpose.cl:15-84 semantics) and a zero fill
 * (used for libxstream_mem_zero so that callers under stream capture get no memset nodes).
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "lxs_kernels.h"

namespace {

struct smm_entry_t {
  int tm, tn, nt, m, n, k, kt;
  const void* fn;
};

struct mfma_entry_t {
  int ti, tj, nwi, kb, m, n, k;
  int ep; /* 1: one wavefront per run (variant 1); > 1: ep wavefronts share each run (variant 2) */
  int b4; /* 1: entry-parallel on 4x4 blocks (variant 3, FP64); 2: the same splitting k (variant 4) */
  const void* fn;
};

const size_t lds_limit = 65536;

/* In-place transpose of listed blocks: column-major m x n -> row-major m x n. Wavefront w of a work-group
 * transposes block blockIdx.x * W + w on its own (W = blockDim.x / 64 blocks in flight per work-group),
 * staging it in its own m n elements of LDS; loads and stores go 4 per lane at a time. */
template <typename T>
__global__ void __launch_bounds__(256) trans_stack(const int* __restrict__ trs, int offset, int stack_size,
  T* __restrict__ data, int m, int n) {
  extern __shared__ double lxs_tbuf[];
  const int mn = m * n;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * (int)(blockDim.x >> 6) + wave;
  T* const buf = reinterpret_cast<T*>(lxs_tbuf) + (size_t)wave * mn;
  T* const mat = data + (i < stack_size ? trs[offset + i] : 0);
  if (i < stack_size) {
    for (int x0 = 0; x0 < mn; x0 += 256) {
      T v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int x = x0 + 64 * u + lane;
        v[u] = (x < mn ? mat[x] : T(0));
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int x = x0 + 64 * u + lane;
        if (x < mn) buf[x] = v[u];
      }
    }
  }
  __syncthreads();
  if (i < stack_size) {
    for (int y = lane; y < mn; y += 64) {
      const int r = y / n, c = y - r * n; /* row-major position (r, c) */
      mat[y] = buf[c * m + r];
    }
  }
}

__global__ void __launch_bounds__(256) zero_fill(unsigned char* __restrict__ p, size_t nbytes) {
  const size_t head = (16 - ((uintptr_t)p & 15)) & 15;
  const size_t pre = (head < nbytes ? head : nbytes);
  const size_t nvec = (nbytes - pre) / 16;
  const size_t tail = pre + nvec * 16;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint4* const v = reinterpret_cast<uint4*>(p + pre);
  for (size_t i = tid; i < nvec; i += stride) v[i] = make_uint4(0, 0, 0, 0);
  if (tid < pre) p[tid] = 0;
  for (size_t i = tail + tid; i < nbytes; i += stride) p[i] = 0;
}

} /* namespace */

extern "C" const void* lxs_smm_table_f64(int* count);
extern "C" const void* lxs_smm_table_f32(int* count);
extern "C" const void* lxs_mfma_table_f64(int* count);
extern "C" const void* lxs_mfma_table_f32(int* count);

static const mfma_entry_t* mfma_table(int dtype, int* count) {
  const void* t = (3 == dtype ? lxs_mfma_table_f64(count) : (1 == dtype ? lxs_mfma_table_f32(count) : NULL));
  if (NULL == t) *count = 0;
  return (const mfma_entry_t*)t;
}

/* threads of an MFMA kernel: nwi wavefronts splitting the tiles, or ep wavefronts sharing each run */
static int mfma_threads(const mfma_entry_t* e) { return 64 * (1 < e->ep ? e->ep : e->nwi); }
static int mfma_variant(const mfma_entry_t* e) { return (0 != e->b4 ? 2 + e->b4 : (1 < e->ep ? 2 : 1)); }
/* LDS bytes of an MFMA kernel: the partial blocks of the ep wavefronts (16 x 16 tiles, or one register per
   4x4-block instruction and lane) */
static size_t mfma_lds(const mfma_entry_t* e, int dtype) {
  return (1 < e->ep ? (size_t)e->ep * e->ti * e->tj * (2 == e->b4 ? 16 : (1 == e->b4 ? 64 : 256)) * (3 == dtype ? 8 : 4) : 0);
}
/* a runtime-shape kernel covers (m, n): 16x16 tiles up to 16 ti x 16 tj, with at least two rows and columns
   where the kernel loads element pairs (an even number of tiles in both dimensions); 4x4 blocks with exactly
   the ti = ceil(m / 8), tj = ceil(n / 8) the shape needs (each lane loads ti consecutive rows, tj columns), or
   ti = ceil(m / 4), tj = ceil(n / 4) when the blocks split k */
static int mfma_fits(const mfma_entry_t* e, int m, int n) {
  if (0 != e->b4) {
    const int r = (2 == e->b4 ? 4 : 8);
    return m <= r * e->ti && m > r * (e->ti - 1) && n <= r * e->tj && n > r * (e->tj - 1);
  }
  if (0 == e->ti % 2 && 0 == e->tj % 2 && (2 > m || 2 > n)) return 0;
  return m <= 16 * e->ti && n <= 16 * e->tj;
}

/* MFMA kernel for (dtype, variant 1 to 4, tiles tm x tn, nt threads, kt = k per stage or 0): exact shape first */
static const mfma_entry_t* mfma_find(int dtype, int variant, int m, int n, int k, int tm, int tn, int nt, int kt) {
  int count = 0, i;
  const mfma_entry_t* const table = mfma_table(dtype, &count);
  const mfma_entry_t* generic = NULL;
  for (i = 0; i < count; ++i) {
    const mfma_entry_t* const e = table + i;
    if (mfma_variant(e) != variant || e->ti != tm || e->tj != tn || mfma_threads(e) != nt ||
        (0 != kt && 4 * e->kb != kt)) {
      continue;
    }
    if (e->m == m && e->n == n && e->k == k) return e;
    if (0 == e->m && NULL == generic && mfma_fits(e, m, n)) generic = e;
  }
  return generic;
}

static const smm_entry_t* smm_table(int dtype, int* count) {
  const void* t = (3 == dtype ? lxs_smm_table_f64(count) : (1 == dtype ? lxs_smm_table_f32(count) : NULL));
  if (NULL == t) *count = 0;
  return (const smm_entry_t*)t;
}

/* best entry for (dtype, tm, tn, nt): a shape-specialised kernel if compiled, else the runtime-shape one */
static const smm_entry_t* smm_find(int dtype, int m, int n, int k, int tm, int tn, int nt) {
  int count = 0, i;
  const smm_entry_t* const table = smm_table(dtype, &count);
  const smm_entry_t* generic = NULL;
  for (i = 0; i < count; ++i) {
    const smm_entry_t* const e = table + i;
    if (e->tm != tm || e->tn != tn || e->nt != nt) continue;
    if (e->m == m && e->n == n && e->k == k) return e;
    if (0 == e->m && NULL == generic) generic = e;
  }
  return generic;
}

extern "C" size_t lxs_lds_limit(void) { return lds_limit; }

static size_t smm_lds(const lxs_smm_args_t* a, int kt) {
  const int tiles_m = (a->m + a->tm - 1) / a->tm, tiles_n = (a->n + a->tn - 1) / a->tn;
  const int ntiles = tiles_m * tiles_n;
  const size_t ts = (3 == a->dtype ? 8 : 4);
  int groups;
  size_t slab, red;
  if (0 >= ntiles || ntiles > a->nt || 0 >= kt) return 0;
  groups = a->nt / ntiles;
  slab = (size_t)kt * (size_t)(tiles_m * a->tm + tiles_n * a->tn);
  red = (1 < groups ? (size_t)(tiles_m * a->tm) * (size_t)(tiles_n * a->tn) : 0);
  return ts * (size_t)groups * (slab > red ? slab : red);
}

/* 1 when the vector-kernel configuration can be launched (its LDS fits; none for wavefront-per-run kernels) */
extern "C" int lxs_smm_launchable(const lxs_smm_args_t* args) {
  const smm_entry_t* e;
  size_t lds;
  int kt;
  if (1 == args->variant) return 1;
  if (2 <= args->variant && args->variant <= 4) {
    const mfma_entry_t* const me = mfma_find(args->dtype, args->variant, args->m, args->n, args->k, args->tm, args->tn,
      args->nt, args->kt);
    return (NULL != me && mfma_lds(me, args->dtype) <= lds_limit) ? 1 : 0;
  }
  e = smm_find(args->dtype, args->m, args->n, args->k, args->tm, args->tn, args->nt);
  if (NULL == e) return 0;
  if (0 > e->kt) return 1;
  kt = (0 < e->kt ? e->kt : (0 < args->kt ? (args->kt < args->k ? args->kt : args->k) : args->k));
  lds = smm_lds(args, kt);
  return (0 < lds && lds <= lds_limit) ? 1 : 0;
}

extern "C" size_t lxs_smm_lds_bytes(const lxs_smm_args_t* args) {
  const smm_entry_t* e;
  int kt;
  if (1 == args->variant) return 0;
  if (2 <= args->variant && args->variant <= 4) {
    const mfma_entry_t* const me = mfma_find(args->dtype, args->variant, args->m, args->n, args->k, args->tm, args->tn,
      args->nt, args->kt);
    return (NULL != me ? mfma_lds(me, args->dtype) : 0);
  }
  e = smm_find(args->dtype, args->m, args->n, args->k, args->tm, args->tn, args->nt);
  if (NULL == e || 0 > e->kt) return 0; /* no LDS for the wavefront-per-run kernels */
  kt = (0 < e->kt ? e->kt : (0 < args->kt ? (args->kt < args->k ? args->kt : args->k) : args->k));
  return smm_lds(args, kt);
}

extern "C" int lxs_smm_available(int dtype, int variant, int m, int n, int k, int tm, int tn, int nt, int kt) {
  if (1 <= variant && variant <= 4) {
    const mfma_entry_t* const e = mfma_find(dtype, variant, m, n, k, tm, tn, nt, kt);
    return (NULL == e ? 0 : (0 != e->m ? 2 : 1));
  }
  if (0 == variant) {
    const smm_entry_t* const e = smm_find(dtype, m, n, k, tm, tn, nt);
    return (NULL == e ? 0 : (0 != e->m ? 2 : 1));
  }
  return 0;
}

extern "C" int lxs_smm_variants(int dtype, int* q, int cap) {
  int count = 0, i, n = 0;
  const smm_entry_t* const table = smm_table(dtype, &count);
  const mfma_entry_t* mtable;
  for (i = 0; i < count; ++i) {
    if (0 != table[i].m) continue;
    if (n < cap && NULL != q) {
      q[5 * n + 0] = 0;
      q[5 * n + 1] = table[i].tm;
      q[5 * n + 2] = table[i].tn;
      q[5 * n + 3] = table[i].nt;
      q[5 * n + 4] = 0;
    }
    ++n;
  }
  mtable = mfma_table(dtype, &count);
  for (i = 0; i < count; ++i) {
    if (0 != mtable[i].m) continue;
    if (n < cap && NULL != q) {
      q[5 * n + 0] = mfma_variant(mtable + i);
      q[5 * n + 1] = mtable[i].ti;
      q[5 * n + 2] = mtable[i].tj;
      q[5 * n + 3] = mfma_threads(mtable + i);
      q[5 * n + 4] = 4 * mtable[i].kb;
    }
    ++n;
  }
  return n;
}

/* every compiled kernel usable for (m, n, k): runtime-shape kernels and kernels compiled for exactly
   this shape, as (variant, tm, tn, nt, kt) quintuples without duplicates */
extern "C" int lxs_smm_candidates(int dtype, int m, int n, int k, int* q, int cap) {
  int count = 0, i, j, nq = 0;
  const smm_entry_t* const table = smm_table(dtype, &count);
  const mfma_entry_t* mtable;
  for (i = 0; i < count + 1; ++i) { /* vector kernels (runtime shape, or compiled for this shape) */
    int v[5];
    if (i == count) break;
    if (0 != table[i].m && (table[i].m != m || table[i].n != n || table[i].k != k)) continue;
    v[0] = 0;
    v[1] = table[i].tm;
    v[2] = table[i].tn;
    v[3] = table[i].nt;
    v[4] = 0;
    for (j = 0; j < nq && j < cap; ++j) {
      if (q[5 * j] == v[0] && q[5 * j + 1] == v[1] && q[5 * j + 2] == v[2] && q[5 * j + 3] == v[3]) break;
    }
    if (j == nq) {
      if (nq < cap) {
        for (j = 0; j < 5; ++j) q[5 * nq + j] = v[j];
      }
      ++nq;
    }
  }
  mtable = mfma_table(dtype, &count);
  for (i = 0; i < count; ++i) {
    const mfma_entry_t* const e = mtable + i;
    int v[5];
    if (0 != e->m ? (e->m != m || e->n != n || e->k != k) : !mfma_fits(e, m, n)) continue;
    if (mfma_lds(e, dtype) > lds_limit) continue;
    v[0] = mfma_variant(e);
    v[1] = e->ti;
    v[2] = e->tj;
    v[3] = mfma_threads(e);
    v[4] = 4 * e->kb;
    for (j = 0; j < nq && j < cap; ++j) {
      if (q[5 * j] == v[0] && q[5 * j + 1] == v[1] && q[5 * j + 2] == v[2] && q[5 * j + 3] == v[3] &&
          q[5 * j + 4] == v[4]) {
        break;
      }
    }
    if (j == nq) {
      if (nq < cap) {
        for (j = 0; j < 5; ++j) q[5 * nq + j] = v[j];
      }
      ++nq;
    }
  }
  return nq;
}

extern "C" int lxs_smm_specialised(int dtype, int variant, int m, int n, int k, int* q) {
  int count = 0, i;
  if (1 <= variant && variant <= 4) {
    const mfma_entry_t* const table = mfma_table(dtype, &count);
    for (i = 0; i < count; ++i) {
      if (table[i].m == m && table[i].n == n && table[i].k == k && mfma_variant(table + i) == variant) {
        q[0] = variant;
        q[1] = table[i].ti;
        q[2] = table[i].tj;
        q[3] = mfma_threads(table + i);
        q[4] = 4 * table[i].kb;
        return i;
      }
    }
  }
  else {
    const smm_entry_t* const table = smm_table(dtype, &count);
    for (i = 0; i < count; ++i) {
      if (table[i].m == m && table[i].n == n && table[i].k == k) {
        q[0] = 0;
        q[1] = table[i].tm;
        q[2] = table[i].tn;
        q[3] = table[i].nt;
        q[4] = table[i].kt;
        return i;
      }
    }
  }
  return -1;
}

extern "C" int lxs_smm_launch(const lxs_smm_args_t* a, void* hip_stream) {
  int kt, bs, atomic, ps, pz, tsplit = a->tsplit, stack_size = a->stack_size, m = a->m, n = a->n, k = a->k;
  size_t lds;
  unsigned int nblocks;
  const void* fn;
  void* c;
  const void *pa, *pb;
  const int* stack;
  void* args[14];
  if (1 <= a->variant && a->variant <= 4) {
    const mfma_entry_t* const e = mfma_find(a->dtype, a->variant, a->m, a->n, a->k, a->tm, a->tn, a->nt, a->kt);
    if (NULL == e) return LXS_ENOVARIANT;
    fn = e->fn;
    kt = 4 * e->kb;
    lds = mfma_lds(e, a->dtype);
    if (lds > lds_limit) return LXS_ELIMIT;
  }
  else {
    const smm_entry_t* const e = smm_find(a->dtype, a->m, a->n, a->k, a->tm, a->tn, a->nt);
    if (NULL == e) return LXS_ENOVARIANT;
    fn = e->fn;
    kt = (0 < e->kt ? e->kt : (0 < a->kt ? (a->kt < a->k ? a->kt : a->k) : a->k));
    lds = (0 > e->kt ? 0 : smm_lds(a, kt)); /* kt < 0: wavefront-per-run kernel without LDS */
    if (0 <= e->kt && (0 == lds || lds > lds_limit)) return LXS_ELIMIT;
  }
  if (0 >= stack_size) return 0;
  bs = (0 < a->bs ? a->bs : 32);
  nblocks = (unsigned int)((stack_size + bs - 1) / bs);
  atomic = a->atomic;
  c = a->c;
  pa = a->a;
  pb = a->b;
  /* record stride and index base; a DBM task record (m, n, k, a, b, c) is read from its fourth int */
  ps = (0 == a->pfmt ? 3 : 6);
  pz = (0 == a->pfmt ? 1 : 0);
  stack = a->stack + (0 == a->pfmt ? 0 : 3);
  args[0] = &c;
  args[1] = &pa;
  args[2] = &pb;
  args[3] = &stack;
  args[4] = &stack_size;
  args[5] = &bs;
  args[6] = &m;
  args[7] = &n;
  args[8] = &k;
  args[9] = &kt;
  args[10] = &atomic;
  args[11] = &ps;
  args[12] = &pz;
  args[13] = &tsplit;
  return (int)hipLaunchKernel(fn, dim3(nblocks), dim3((unsigned int)a->nt), args, lds, (hipStream_t)hip_stream);
}

extern "C" int lxs_trans_launch(const int* trs_stack, int offset, int stack_size, void* data, int dtype, int m, int n,
  void* hip_stream) {
  const size_t ts = (3 == dtype ? 8 : 4);
  const size_t blk = ts * (size_t)m * (size_t)n; /* LDS bytes per block (one wavefront each) */
  unsigned int waves, nblocks;
  if (0 >= stack_size) return 0;
  if (blk > lds_limit) return LXS_ELIMIT;
  waves = (unsigned int)(lds_limit / blk < 4 ? lds_limit / blk : 4);
  nblocks = (unsigned int)((stack_size + (int)waves - 1) / (int)waves);
  if (3 == dtype) {
    trans_stack<double><<<nblocks, 64 * waves, waves * blk, (hipStream_t)hip_stream>>>(trs_stack, offset, stack_size,
      (double*)data, m, n);
  }
  else if (1 == dtype) {
    trans_stack<float><<<nblocks, 64 * waves, waves * blk, (hipStream_t)hip_stream>>>(trs_stack, offset, stack_size,
      (float*)data, m, n);
  }
  else {
    return LXS_ENOVARIANT;
  }
  return (int)hipGetLastError();
}

extern "C" int lxs_zero_launch(void* ptr, size_t nbytes, void* hip_stream) {
  size_t nblocks;
  if (0 == nbytes) return 0;
  nblocks = (nbytes / 16 + 255) / 256;
  if (nblocks < 1) nblocks = 1;
  if (nblocks > 4096) nblocks = 4096;
  zero_fill<<<(unsigned int)nblocks, 256, 0, (hipStream_t)hip_stream>>>((unsigned char*)ptr, nbytes);
  return (int)hipGetLastError();
}
