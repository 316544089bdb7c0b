/*
 * Ozaki host layer: context, scheme selection, device scratch, K passes and the complex 3M method.
 *
 * Entry points and argument meaning follow samples/ozaki/ozaki_opencl.h:328-440; settings are
 * resolved as in samples/ozaki/ozaki_opencl.c:348-475 (defaults fp64 8 slices / 16 primes, fp32
 * 4 / 9; Scheme 1 cutoff 2(nslices-1) - oztrim; Scheme 2 drops 2 oztrim mantissa bits) and the
 * adaptive choice as in samples/ozaki/ozaki_gemm.c:520-587 (Scheme 1 while its pair count is below
 * the prime count plus the reconstruction's share). The device work is ozaki.hip: per-operand
 * preprocessing into int8 planes (digits or residues), one int8 GEMM launch per K pass, then the FP64
 * recombination (Scheme 1) or the CRT reconstruction (Scheme 2).
 */
#include "lxs_internal.h"
#include "../../include/ozaki.h"
#include <string.h>

#define OZ_ALIGN 256
#define OZ_TILE_MN 256    /* output tile of the int8 GEMM; row padding of the planes */
#define OZ_TILE_K 128     /* k padding of the planes */
#define OZ_UNIT_K 64      /* k per unit of the int8 GEMM */
#define OZ_FLUSH_MAX 2080 /* units per Scheme 1 work item: 2080 * 64 * 127^2 < 2^31 (exact int32 partials) */
#define OZ_FLUSH_CRT 2047 /* units per Scheme 2 work item: 2047 * 64 * 128^2 < 2^31 */
#define OZ_WG_TARGET 1024 /* work-groups the item plan aims for */
#define OZ_KPASS_MAX ((1 << 22) - OZ_TILE_K) /* K per pass (32-bit operand offsets of the int8 GEMM) */
#define OZ_NOALLOC_ITEMS ((int)(LXS_OZ_STATIC_BYTES / ((size_t)OZ_TILE_MN * OZ_TILE_MN)))
#define OZ_XOVER 16.0     /* adaptive choice: reconstruction cost per prime^2, in int8 GEMMs of K = 1 */

typedef unsigned __int128 oz_u128_t;

typedef struct oz_impl_t {
  int* items; /* work items {level, u0, u1, pad} of the current call */
  int items_cap;
  void* arena; /* internal scratch, grown on demand */
  size_t arena_size;
  void* user; /* caller-owned scratch (ozaki_scratch_set) */
  size_t user_size;
  void* carena; /* complex 3M operands and products */
  size_t carena_size;
  void* sarena; /* stack path: digit planes and exponents of the pools */
  size_t sarena_size;
} oz_impl_t;

static const int oz_moduli[LXS_OZ_NPRIMES_MAX] = {LXS_OZ_MODULI};

static size_t oz_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

static int oz_env_int(const char* name, int fallback) {
  const char* const env = getenv(name);
  return (NULL != env && '\0' != *env) ? atoi(env) : fallback;
}

static int oz_mant_bits(const ozaki_context_t* ctx) { return 0 != ctx->use_double ? 53 : 24; }

/* mantissa bits Scheme 2 drops: 2 per trim level, kind 2 only (samples/ozaki/ozaki_opencl.c:428-431) */
static int oz_crt_trim(const ozaki_context_t* ctx) {
  const int levels = (oz_mant_bits(ctx) - 1) / 2;
  return (2 == ctx->kind ? 2 * (ctx->oztrim < levels ? ctx->oztrim : levels) : 0);
}

/* product of the first np moduli */
static oz_u128_t oz_crt_product(int np) {
  oz_u128_t m = 1;
  int p;
  for (p = 0; p < np; ++p) m *= (oz_u128_t)oz_moduli[p];
  return m;
}

/* the integer product of a K pass is exact under np primes when 2 kk (2^bits - 1)^2 < M (|x| < M / 2) */
static int oz_crt_exact(int np, int bits, long kk) {
  const oz_u128_t amax = ((oz_u128_t)1 << bits) - 1;
  return (oz_crt_product(np) - 1) / (2 * amax * amax) >= (oz_u128_t)kk;
}

/* fewest primes (at most npmax) keeping a K pass of kk exact, 0 if npmax do not */
static int oz_crt_primes(int npmax, int bits, long kk) {
  int np;
  for (np = 1; np <= npmax; ++np) {
    if (oz_crt_exact(np, bits, kk)) return np;
  }
  return 0;
}

/* longest K pass npmax primes keep exact (at most 2^30) */
static long oz_crt_kmax(int npmax, int bits) {
  const oz_u128_t amax = ((oz_u128_t)1 << bits) - 1;
  const oz_u128_t k = (oz_crt_product(npmax) - 1) / (2 * amax * amax);
  return (k < ((oz_u128_t)1 << 30) ? (long)k : (1L << 30));
}

LIBXSTREAM_API int ozaki_init(ozaki_context_t* ctx, int tm, int tn, int use_double, int kind, int verbosity, int ndecomp,
  int ozflags, int oztrim, int ozgroups, int maxk) {
  oz_impl_t* impl;
  int nslices, nprimes;
  (void)ozgroups;
  if (NULL == ctx) return EXIT_FAILURE;
  memset(ctx, 0, sizeof(*ctx));
  if (0 >= kind || 3 < kind) kind = 3; /* no request: adaptive */
  if ((0 != tm && OZ_TILE_MN != tm) || (0 != tn && OZ_TILE_MN != tn)) {
    lxs_error(EXIT_FAILURE, "ozaki_init: the output tile is 256 x 256");
    return EXIT_FAILURE;
  }
  nslices = (0 != use_double ? 8 : 4);
  nprimes = (0 != use_double ? 16 : 9);
  if (0 < ndecomp) { /* the slice count for kind 1, else the prime count */
    if (1 == kind) nslices = ndecomp;
    else nprimes = ndecomp;
  }
  if (8 < nslices) nslices = 8; /* 8 digits of 7 bits hold a full FP64 mantissa */
  if (LXS_OZ_NPRIMES_MAX < nprimes) nprimes = LXS_OZ_NPRIMES_MAX; /* 16 moduli: 125 bits */
  if (2 > nprimes) nprimes = 2;
  if (0 > ozflags) ozflags = OZAKI_TRIANGULAR | OZAKI_SYMMETRIZE;
  if (2 != kind && (1 == ozflags || 2 == ozflags)) { /* Scheme 1 pair traversal */
    lxs_error(EXIT_FAILURE, "ozaki_init: ozflags 1 and 2 (symmetric operands only) are not supported, use 0 or 3");
    return EXIT_FAILURE;
  }
  if (0 > oztrim) oztrim = 0;
  if (2 != kind && oztrim > 2 * (nslices - 1)) oztrim = 2 * (nslices - 1);
  impl = (oz_impl_t*)calloc(1, sizeof(oz_impl_t));
  if (NULL == impl) return EXIT_FAILURE;
  /* work items for any call that fits the static scratch of ozaki_gemm_noalloc (an item's output is at
     least one byte per element of a 256 x 256 tile), so that entry point never grows the array */
  impl->items = (int*)malloc(sizeof(int) * 4 * OZ_NOALLOC_ITEMS);
  if (NULL == impl->items) {
    free(impl);
    return EXIT_FAILURE;
  }
  impl->items_cap = OZ_NOALLOC_ITEMS;
  if (0 <= lxs_cfg.device && 0 == lxs_ensure_device()) { /* the static scratch's address on the active device */
    void* scratch = NULL;
    if (0 != lxs_oz_static_scratch(&scratch)) {
      free(impl->items);
      free(impl);
      return lxs_error(EXIT_FAILURE, "ozaki_init: device address of the static scratch");
    }
  }
  ctx->use_double = (0 != use_double ? 1 : 0);
  ctx->kind = kind;
  ctx->nslices = nslices;
  ctx->nprimes = nprimes;
  ctx->ndecomp = (1 == kind ? nslices : nprimes);
  ctx->ozflags = ozflags;
  ctx->oztrim = oztrim;
  ctx->verbosity = verbosity;
  ctx->tm = ctx->tn = OZ_TILE_MN;
  ctx->maxk = (0 < maxk ? maxk : 0);
  ctx->flush = oz_env_int("LIBXSTREAM_OZAKI_FLUSH", OZ_FLUSH_MAX);
  if (1 > ctx->flush || OZ_FLUSH_MAX < ctx->flush) ctx->flush = OZ_FLUSH_MAX;
  ctx->impl = impl;
  if (0 != verbosity) {
    const int cutoff = 2 * (nslices - 1) - (2 == kind ? 0 : oztrim);
    fprintf(stderr, "INFO OZAKI: kind=%d %s, scheme 1: nslices=%d cutoff=%d pairs=%d; scheme 2: primes<=%d trim=%d bits; tile=%dx%d\n",
      kind, ctx->use_double ? "fp64" : "fp32", nslices, cutoff, ozaki_count_pairs(nslices, cutoff, 0), nprimes,
      oz_crt_trim(ctx), OZ_TILE_MN, OZ_TILE_MN);
  }
  return EXIT_SUCCESS;
}

LIBXSTREAM_API void ozaki_destroy(ozaki_context_t* ctx) {
  if (NULL != ctx && NULL != ctx->impl) {
    oz_impl_t* const impl = (oz_impl_t*)ctx->impl;
    if (NULL != impl->arena || NULL != impl->carena || NULL != impl->sarena) {
      (void)hipDeviceSynchronize();
      if (NULL != impl->arena) (void)hipFree(impl->arena);
      if (NULL != impl->carena) (void)hipFree(impl->carena);
      if (NULL != impl->sarena) (void)hipFree(impl->sarena);
    }
    free(impl->items);
    free(impl);
    ctx->impl = NULL;
  }
}

/* rows and columns of a stored operand: op = 'N' keeps (rows, cols) of op(X), 'T'/'C' swaps them */
static int oz_trans(char t) {
  switch (t) {
    case 'N': case 'n': return 0;
    case 'T': case 't': return 1;
    case 'C': case 'c': return 2;
    default: return -1;
  }
}

/* Scheme 1 level cutoff (kind 2 keeps oztrim for the mantissa) */
static int oz_cutoff(const ozaki_context_t* ctx) {
  return 2 * (ctx->nslices - 1) - (2 == ctx->kind ? 0 : ctx->oztrim);
}

/*
 * The scheme of a call (1 or 2, 0 when kind 2 has too few primes for even K = 1) and, for Scheme 2, its
 * prime count and K pass: kind 1 and 2 force their scheme; kind 3 takes Scheme 1 while
 * pairs < np + OZ_XOVER np^2 / K (int8 GEMMs of Scheme 1 against those of Scheme 2 plus its
 * reconstruction, per unit of K).
 */
static int oz_scheme(const ozaki_context_t* ctx, int K, int* np, int* kpass) {
  const int bits = oz_mant_bits(ctx) - oz_crt_trim(ctx);
  long kp = oz_crt_kmax(ctx->nprimes, bits);
  int scheme;
  if (0 < ctx->maxk && ctx->maxk < kp) kp = ctx->maxk;
  if (OZ_KPASS_MAX < kp) kp = OZ_KPASS_MAX;
  if (K < kp) kp = (0 < K ? K : 1);
  *np = oz_crt_primes(ctx->nprimes, bits, kp);
  *kpass = (int)kp;
  if (1 == ctx->kind || (0 == *np && 2 != ctx->kind)) scheme = 1;
  else if (2 == ctx->kind) scheme = (0 != *np ? 2 : 0);
  else {
    const int pairs = ozaki_count_pairs(ctx->nslices, oz_cutoff(ctx), 0);
    scheme = ((double)pairs < *np + OZ_XOVER * *np * *np / (0 < K ? K : 1)) ? 1 : 2;
  }
  if (1 == scheme) {
    *kpass = (0 < ctx->maxk && ctx->maxk < K) ? ctx->maxk : (0 < K ? K : 1);
    if (OZ_KPASS_MAX < *kpass) *kpass = OZ_KPASS_MAX;
  }
  return scheme;
}

/*
 * Work items of the int8 GEMM. Scheme 1: every level L = 0 .. cutoff has npairs(L) * nkb units (pair,
 * k-block of 64); Scheme 2: every prime has nkb units. Levels are cut into items of at most cu units, cu
 * chosen so that items x tiles is about OZ_WG_TARGET work-groups (at least 16 units per item, at most
 * ctx->flush, and 2047 for Scheme 2). zb[L] = first item of level L. Returns the number of items; fills
 * items (4 ints each) when not NULL.
 */
static int oz_plan(const ozaki_context_t* ctx, int crt, int np, int nkb, int tiles, int* items, int* zb) {
  const int s = ctx->nslices, cutoff = (0 != crt ? np - 1 : oz_cutoff(ctx));
  const int cmax = (0 != crt && OZ_FLUSH_CRT < ctx->flush ? OZ_FLUSH_CRT : ctx->flush);
  long total = 0;
  int L, n = 0, cu;
  for (L = 0; L <= cutoff; ++L) {
    const int plo = (L - s + 1 > 0 ? L - s + 1 : 0), phi = (L < s - 1 ? L : s - 1);
    total += (long)(0 != crt ? 1 : phi - plo + 1) * nkb;
  }
  cu = (int)((total * tiles + OZ_WG_TARGET - 1) / OZ_WG_TARGET);
  if (cu < 16) cu = 16;
  if (cu > cmax) cu = cmax;
  for (L = 0; L <= cutoff; ++L) {
    const int plo = (L - s + 1 > 0 ? L - s + 1 : 0), phi = (L < s - 1 ? L : s - 1);
    const int units = (0 != crt ? 1 : phi - plo + 1) * nkb;
    int u;
    if (NULL != zb) zb[L] = n;
    for (u = 0; u < units; u += cu, ++n) {
      if (NULL != items) {
        items[4 * n + 0] = L;
        items[4 * n + 1] = u;
        items[4 * n + 2] = (units - u < cu ? units : u + cu);
        items[4 * n + 3] = 0;
      }
    }
  }
  if (NULL != zb) zb[cutoff + 1] = n;
  return n;
}

typedef struct oz_layout_t {
  int Mp, Np, Kp, kpass, nitems, crt, np, scheme;
  size_t planes_a, planes_b, ea, eb, occ, out, a, b, c, total; /* byte offsets and total */
} oz_layout_t;

static void oz_layout(const ozaki_context_t* ctx, int crt, int np, int kpass, int ta, int tb, int M, int N, int K,
  int lda, int ldb, int ldc, int dev, oz_layout_t* l) {
  const size_t ts = (0 != ctx->use_double ? 8 : 4);
  const int nplanes = (0 != crt ? np : ctx->nslices);
  size_t off = 0;
  l->crt = crt;
  l->np = np;
  l->Mp = (int)oz_up((size_t)(0 < M ? M : 1), OZ_TILE_MN);
  l->Np = (int)oz_up((size_t)(0 < N ? N : 1), OZ_TILE_MN);
  l->Kp = (int)oz_up((size_t)(0 < kpass ? kpass : 1), OZ_TILE_K);
  l->kpass = kpass;
  l->nitems = oz_plan(ctx, crt, np, l->Kp / OZ_UNIT_K, (l->Mp / OZ_TILE_MN) * (l->Np / OZ_TILE_MN), NULL, NULL);
  l->planes_a = off;
  off += oz_up((size_t)nplanes * l->Mp * l->Kp, OZ_ALIGN);
  l->planes_b = off;
  off += oz_up((size_t)nplanes * l->Np * l->Kp, OZ_ALIGN);
  l->ea = off;
  off += oz_up(sizeof(int) * (size_t)l->Mp, OZ_ALIGN);
  l->eb = off;
  off += oz_up(sizeof(int) * (size_t)l->Np, OZ_ALIGN);
  l->occ = off;
  off += OZ_ALIGN; /* 2 x 8 ints */
  l->out = off; /* int32 partials (Scheme 1) or residue bytes (Scheme 2) per item */
  off += oz_up((0 != crt ? 1 : sizeof(int)) * (size_t)l->nitems * l->Mp * l->Np, OZ_ALIGN);
  l->a = l->b = l->c = 0;
  if (0 == dev) {
    l->a = off;
    off += oz_up(ts * (size_t)lda * (size_t)(0 == ta ? K : M), OZ_ALIGN);
    l->b = off;
    off += oz_up(ts * (size_t)ldb * (size_t)(0 == tb ? N : K), OZ_ALIGN);
    l->c = off;
    off += oz_up(ts * (size_t)ldc * (size_t)N, OZ_ALIGN);
  }
  l->total = off;
}

/* layout of a call: the scheme oz_scheme picks for it */
static void oz_layout_call(const ozaki_context_t* ctx, int ta, int tb, int M, int N, int K, int lda, int ldb, int ldc,
  int dev, oz_layout_t* l) {
  int np, kpass;
  const int scheme = oz_scheme(ctx, K, &np, &kpass);
  oz_layout(ctx, 2 == scheme, np, kpass, ta, tb, M, N, K, lda, ldb, ldc, dev, l);
  l->scheme = scheme;
}

LIBXSTREAM_API size_t ozaki_scratch_size(const ozaki_context_t* ctx, char transa, char transb, int M, int N, int K,
  int lda, int ldb, int ldc) {
  oz_layout_t l;
  const int ta = oz_trans(transa), tb = oz_trans(transb);
  if (NULL == ctx || 0 > ta || 0 > tb || 0 > M || 0 > N || 0 > K) return 0;
  oz_layout_call(ctx, ta, tb, M, N, K, lda, ldb, ldc, 0 /*host operands: upper bound*/, &l);
  return l.total;
}

LIBXSTREAM_API int ozaki_scratch_set(ozaki_context_t* ctx, void* dev_mem, size_t nbytes) {
  oz_impl_t* impl;
  if (NULL == ctx || NULL == ctx->impl) return EXIT_FAILURE;
  impl = (oz_impl_t*)ctx->impl;
  impl->user = dev_mem;
  impl->user_size = (NULL != dev_mem ? nbytes : 0);
  if (NULL == dev_mem && NULL != impl->arena) { /* revert to the internal arena: release it */
    (void)hipDeviceSynchronize();
    (void)hipFree(impl->arena);
    impl->arena = NULL;
    impl->arena_size = 0;
  }
  return EXIT_SUCCESS;
}

LIBXSTREAM_API void* ozaki_scratch_get(const ozaki_context_t* ctx, size_t* nbytes) {
  const oz_impl_t* impl = (NULL != ctx ? (const oz_impl_t*)ctx->impl : NULL);
  void* ptr = NULL;
  size_t size = 0;
  if (NULL != impl) {
    if (NULL != impl->user) {
      ptr = impl->user;
      size = impl->user_size;
    }
    else {
      ptr = impl->arena;
      size = impl->arena_size;
    }
  }
  if (NULL != nbytes) *nbytes = size;
  return ptr;
}

LIBXSTREAM_API ozaki_tile_t ozaki_tile_select(const ozaki_context_t* ctx, int M, int N, int rtm, int rtn) {
  ozaki_tile_t tile;
  (void)ctx, (void)M, (void)N, (void)rtm, (void)rtn;
  tile.m = tile.n = OZ_TILE_MN;
  return tile;
}

LIBXSTREAM_API int ozaki_npanel(const ozaki_context_t* ctx, int M, int N, int tm, int tn) {
  (void)ctx, (void)M, (void)tm, (void)tn;
  return N;
}

LIBXSTREAM_API void ozaki_invalidate_cache(ozaki_context_t* ctx, const void* a, const void* b) {
  (void)ctx, (void)a, (void)b;
}

/* device memory of at least nbytes: the caller's scratch when large enough, else the internal arena */
static char* oz_scratch(oz_impl_t* impl, size_t nbytes, hipStream_t hs) {
  if (NULL != impl->user && impl->user_size >= nbytes) return (char*)impl->user;
  if (impl->arena_size < nbytes) {
    if (NULL != impl->arena) {
      if (0 != lxs_hip(hipStreamSynchronize(hs), "ozaki_gemm: stream synchronize")) return NULL;
      (void)hipFree(impl->arena);
      impl->arena = NULL;
      impl->arena_size = 0;
    }
    if (0 != lxs_hip(hipMalloc(&impl->arena, nbytes), "ozaki_gemm: scratch allocation")) {
      impl->arena = NULL;
      return NULL;
    }
    impl->arena_size = nbytes;
  }
  return (char*)impl->arena;
}

/* fixed != NULL: the call's scratch is fixed[0 .. fixed_size) (ozaki_gemm_noalloc): nothing is allocated,
   device operands only, no profile records; a call that needs more fails */
static int oz_gemm_real(ozaki_context_t* ctx, hipStream_t hs, int ta, int tb, int M, int N, int K, double alpha,
  const void* a, int lda, const void* b, int ldb, double beta, void* c, int ldc, int dev, char* fixed,
  size_t fixed_size) {
  oz_impl_t* const impl = (oz_impl_t*)ctx->impl;
  const int dtype = (0 != ctx->use_double ? 3 : 1);
  const size_t ts = (0 != ctx->use_double ? 8 : 4);
  const int arows = (0 == ta ? M : K), acols = (0 == ta ? K : M);
  const int brows = (0 == tb ? K : N), bcols = (0 == tb ? N : K);
  const int s = ctx->nslices, cutoff = oz_cutoff(ctx);
  oz_layout_t l;
  char* base;
  const char *da = (const char*)a, *db = (const char*)b;
  char* dc = (char*)c;
  int result = 0, k0;
  if (lda < (1 > arows ? 1 : arows) || ldb < (1 > brows ? 1 : brows) || ldc < (1 > M ? 1 : M)) {
    return lxs_error(EXIT_FAILURE, "ozaki_gemm: leading dimension too small");
  }
  if (0 == M || 0 == N) return EXIT_SUCCESS;
  if ((0 != K && (NULL == a || NULL == b)) || NULL == c) return lxs_error(EXIT_FAILURE, "ozaki_gemm: NULL operand");
  oz_layout_call(ctx, ta, tb, M, N, K, lda, ldb, ldc, dev, &l);
  if (0 == l.scheme) return lxs_error(EXIT_FAILURE, "ozaki_gemm: too few CRT primes for an exact product");
  if (NULL != fixed) {
    if (0 == dev) return lxs_error(EXIT_FAILURE, "ozaki_gemm_noalloc: operands must be device pointers");
    if (l.total > fixed_size || l.nitems > impl->items_cap) {
      return lxs_error(EXIT_FAILURE, "ozaki_gemm_noalloc: the call needs more than the static scratch "
                                     "(LXS_OZ_STATIC_BYTES)");
    }
    base = fixed;
  }
  else {
    base = oz_scratch(impl, l.total, hs);
    if (NULL == base) return EXIT_FAILURE;
  }
  if (0 == dev) { /* host operands: copy in */
    da = base + l.a;
    db = base + l.b;
    dc = base + l.c;
    if (0 != K) {
      result |= lxs_hip(hipMemcpy2DAsync(base + l.a, ts * lda, a, ts * lda, ts * arows, acols, hipMemcpyHostToDevice, hs),
        "ozaki_gemm: copy A");
      result |= lxs_hip(hipMemcpy2DAsync(base + l.b, ts * ldb, b, ts * ldb, ts * brows, bcols, hipMemcpyHostToDevice, hs),
        "ozaki_gemm: copy B");
    }
    if (0 != beta) {
      result |= lxs_hip(hipMemcpy2DAsync(dc, ts * ldc, c, ts * ldc, ts * M, N, hipMemcpyHostToDevice, hs), "ozaki_gemm: copy C");
    }
    if (0 != result) return EXIT_FAILURE;
  }
  if (0 == K || 0 == alpha) {
    result = lxs_oz_scale(dtype, dc, ldc, M, N, beta, hs);
  }
  else {
    /* op(A)(i, k) at da[i * ars + k * aks]; op(B)(k, j) at db[k * bks + j * bcs] */
    const long ars = (0 == ta ? 1 : lda), aks = (0 == ta ? lda : 1);
    const long bks = (0 == tb ? 1 : ldb), bcs = (0 == tb ? ldb : 1);
    const int tiles = (l.Mp / OZ_TILE_MN) * (l.Np / OZ_TILE_MN);
    const int trim = oz_crt_trim(ctx), mb = oz_mant_bits(ctx);
    int* const occ = (int*)(base + l.occ);
    int zb[LXS_OZ_NPRIMES_MAX + 1];
    if (impl->items_cap < l.nitems) {
      int* const items = (int*)realloc(impl->items, sizeof(int) * 4 * (size_t)l.nitems);
      if (NULL == items) return lxs_error(EXIT_FAILURE, "ozaki_gemm: out of memory");
      impl->items = items;
      impl->items_cap = l.nitems;
    }
    if (0 != ctx->verbosity && 1 < ctx->verbosity) {
      fprintf(stderr, "INFO OZAKI: M=%d N=%d K=%d scheme %d, %d %s, K pass %d\n", M, N, K, 0 != l.crt ? 2 : 1,
        0 != l.crt ? l.np : ozaki_count_pairs(s, cutoff, 0), 0 != l.crt ? "primes" : "pairs", l.kpass);
    }
    for (k0 = 0; k0 < K && 0 == result; k0 += l.kpass) {
      const int kk = (K - k0 < l.kpass ? K - k0 : l.kpass);
      const int Kp = (int)oz_up((size_t)kk, OZ_TILE_K);
      /* the last pass may need fewer primes */
      const int np = (0 != l.crt ? oz_crt_primes(l.np, mb - trim, kk) : 0);
      const int nitems = oz_plan(ctx, l.crt, np, Kp / OZ_UNIT_K, tiles, impl->items, zb);
      const double nprod = (0 != l.crt ? np : ozaki_count_pairs(s, cutoff, 0));
      const int nplanes = (0 != l.crt ? np : s);
      void* token = (NULL == fixed ? lxs_profile_begin(hs) : NULL);
      if (0 != l.crt) {
        result = lxs_oz_crt_preprocess(dtype, da + ts * (size_t)k0 * aks, ars, aks, M, kk, np, 1, trim, l.Mp, Kp,
          (int*)(base + l.ea), (signed char*)(base + l.planes_a), hs);
        if (0 == result) {
          result = lxs_oz_crt_preprocess(dtype, db + ts * (size_t)k0 * bks, bcs, bks, N, kk, np, 0, trim, l.Np, Kp,
            (int*)(base + l.eb), (signed char*)(base + l.planes_b), hs);
        }
      }
      else {
        result = lxs_oz_preprocess(dtype, da + ts * (size_t)k0 * aks, ars, aks, M, kk, s, l.Mp, Kp, (int*)(base + l.ea), occ,
          (signed char*)(base + l.planes_a), hs);
        if (0 == result) {
          result = lxs_oz_preprocess(dtype, db + ts * (size_t)k0 * bks, bcs, bks, N, kk, s, l.Np, Kp, (int*)(base + l.eb),
            occ + 8, (signed char*)(base + l.planes_b), hs);
        }
      }
      lxs_profile_end(token, hs, "ozaki_pre", 0.0, (double)ts * ((double)M + N) * kk + (double)nplanes * (l.Mp + l.Np) * Kp);
      if (0 == result) { /* int8 work of every pair up to the cutoff (pairs of all-zero planes are skipped), or every prime */
        token = (NULL == fixed ? lxs_profile_begin(hs) : NULL);
        result = lxs_oz_level_gemm(l.crt, (const signed char*)(base + l.planes_a), (const signed char*)(base + l.planes_b),
          occ, occ + 8, impl->items, nitems, nplanes, l.Mp, l.Np, Kp, base + l.out, hs);
        lxs_profile_end(token, hs, "ozaki_i8", 2.0 * nprod * l.Mp * l.Np * Kp,
          (double)nplanes * (l.Mp + l.Np) * Kp + (0 != l.crt ? 1.0 : 4.0) * nitems * l.Mp * l.Np);
      }
      if (0 == result) {
        const double bc = (double)ts * M * N * (0 == k0 && 0 == beta ? 1 : 2);
        token = (NULL == fixed ? lxs_profile_begin(hs) : NULL);
        if (0 != l.crt) {
          result = lxs_oz_crt_recombine(dtype, (const unsigned char*)(base + l.out), zb, np, l.Mp, l.Np,
            (const int*)(base + l.ea), (const int*)(base + l.eb), M, N, 2 * trim - 2 * mb, alpha, (0 == k0 ? beta : 1.0), dc,
            ldc, hs);
          lxs_profile_end(token, hs, "ozaki_fp64", 0.0, (double)nitems * M * N + bc);
        }
        else {
          result = lxs_oz_recombine(dtype, (const int*)(base + l.out), zb, cutoff, l.Mp, l.Np, (const int*)(base + l.ea),
            (const int*)(base + l.eb), M, N, alpha, (0 == k0 ? beta : 1.0), dc, ldc, hs);
          lxs_profile_end(token, hs, "ozaki_fp64", 0.0, 4.0 * nitems * M * N + bc);
        }
      }
    }
  }
  if (0 != result) return lxs_error(EXIT_FAILURE, "ozaki_gemm: kernel launch failed");
  if (0 == dev) {
    result = lxs_hip(hipMemcpy2DAsync(c, ts * ldc, dc, ts * ldc, ts * M, N, hipMemcpyDeviceToHost, hs), "ozaki_gemm: copy C back");
  }
  return (0 == result ? EXIT_SUCCESS : EXIT_FAILURE);
}

LIBXSTREAM_API int ozaki_gemm(ozaki_context_t* ctx, libxstream_stream_t* stream, char transa, char transb, int M, int N,
  int K, double alpha, const void* a, int lda, const void* b, int ldb, double beta, void* c, int ldc, int dev) {
  const int ta = oz_trans(transa), tb = oz_trans(transb);
  hipStream_t hs;
  void* token;
  int result;
  if (NULL == ctx || NULL == ctx->impl) return lxs_error(EXIT_FAILURE, "ozaki_gemm: context not initialised");
  if (0 > ta || 0 > tb || 0 > M || 0 > N || 0 > K) return lxs_error(EXIT_FAILURE, "ozaki_gemm: invalid argument");
  if (0 != lxs_ensure_device()) return EXIT_FAILURE;
  hs = lxs_stream_hip(stream);
  token = lxs_profile_begin(hs);
  result = oz_gemm_real(ctx, hs, ta, tb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, dev, NULL, 0);
  lxs_profile_end(token, hs, "ozaki", 2.0 * M * N * K, (double)(0 != ctx->use_double ? 8 : 4) * ((double)M * K + (double)K * N + 2.0 * M * N));
  return result;
}

LIBXSTREAM_API int ozaki_gemm_noalloc(ozaki_context_t* ctx, libxstream_stream_t* stream, char transa, char transb,
  int M, int N, int K, double alpha, const void* a, int lda, const void* b, int ldb, double beta, void* c, int ldc) {
  const int ta = oz_trans(transa), tb = oz_trans(transb);
  void* scratch = NULL;
  if (NULL == ctx || NULL == ctx->impl) return lxs_error(EXIT_FAILURE, "ozaki_gemm_noalloc: context not initialised");
  if (0 > ta || 0 > tb || 0 > M || 0 > N || 0 > K) return lxs_error(EXIT_FAILURE, "ozaki_gemm_noalloc: invalid argument");
  if (0 != lxs_ensure_device()) return EXIT_FAILURE;
  if (0 != lxs_oz_static_scratch(&scratch)) {
    return lxs_error(EXIT_FAILURE, "ozaki_gemm_noalloc: device address of the static scratch");
  }
  return oz_gemm_real(ctx, lxs_stream_hip(stream), ta, tb, M, N, K, alpha, a, lda, b, ldb, beta, c, ldc, 1,
    (char*)scratch, LXS_OZ_STATIC_BYTES);
}

LIBXSTREAM_API int ozaki_gemm_complex(ozaki_context_t* ctx, libxstream_stream_t* stream, char transa, char transb, int M,
  int N, int K, const double* alpha, const void* a, int lda, const void* b, int ldb, const double* beta, void* c, int ldc) {
  const int ta = oz_trans(transa), tb = oz_trans(transb);
  const int dtype = (NULL != ctx && 0 != ctx->use_double ? 3 : 1);
  const size_t ts = (3 == dtype ? 8 : 4);
  oz_impl_t* impl;
  hipStream_t hs;
  int arows, acols, brows, bcols, result = 0;
  size_t na, nb, nc, need;
  char* base;
  if (NULL == ctx || NULL == ctx->impl || NULL == alpha || NULL == beta) {
    return lxs_error(EXIT_FAILURE, "ozaki_gemm_complex: invalid argument");
  }
  if (0 > ta || 0 > tb || 0 > M || 0 > N || 0 > K) return lxs_error(EXIT_FAILURE, "ozaki_gemm_complex: invalid argument");
  arows = (0 == ta ? M : K), acols = (0 == ta ? K : M);
  brows = (0 == tb ? K : N), bcols = (0 == tb ? N : K);
  if (lda < (1 > arows ? 1 : arows) || ldb < (1 > brows ? 1 : brows) || ldc < (1 > M ? 1 : M)) {
    return lxs_error(EXIT_FAILURE, "ozaki_gemm_complex: leading dimension too small");
  }
  if (0 == M || 0 == N) return EXIT_SUCCESS;
  if (0 != lxs_ensure_device()) return EXIT_FAILURE;
  impl = (oz_impl_t*)ctx->impl;
  hs = lxs_stream_hip(stream);
  na = oz_up(ts * (size_t)arows * acols, OZ_ALIGN);
  nb = oz_up(ts * (size_t)brows * bcols, OZ_ALIGN);
  nc = oz_up(ts * (size_t)M * N, OZ_ALIGN);
  need = 3 * (na + nb + nc);
  if (impl->carena_size < need) {
    if (NULL != impl->carena) {
      if (0 != lxs_hip(hipStreamSynchronize(hs), "ozaki_gemm_complex: stream synchronize")) return EXIT_FAILURE;
      (void)hipFree(impl->carena);
      impl->carena = NULL;
      impl->carena_size = 0;
    }
    if (0 != lxs_hip(hipMalloc(&impl->carena, need), "ozaki_gemm_complex: scratch allocation")) return EXIT_FAILURE;
    impl->carena_size = need;
  }
  base = (char*)impl->carena;
  {
    char *const ar = base, *const ai = base + na, *const as = base + 2 * na;
    char *const br = base + 3 * na, *const bi = br + nb, *const bs = br + 2 * nb;
    char *const t1 = base + 3 * (na + nb), *const t2 = t1 + nc, *const t3 = t1 + 2 * nc;
    const char tra = (0 == ta ? 'N' : 'T'), trb = (0 == tb ? 'N' : 'T');
    if (0 < K) {
      result |= lxs_oz_csplit(dtype, a, lda, arows, acols, 2 == ta, ar, ai, as, hs);
      result |= lxs_oz_csplit(dtype, b, ldb, brows, bcols, 2 == tb, br, bi, bs, hs);
    }
    if (0 != result) return lxs_error(EXIT_FAILURE, "ozaki_gemm_complex: kernel launch failed");
    if (EXIT_SUCCESS != ozaki_gemm(ctx, stream, tra, trb, M, N, K, 1.0, ar, arows, br, brows, 0.0, t1, M, 1) ||
        EXIT_SUCCESS != ozaki_gemm(ctx, stream, tra, trb, M, N, K, 1.0, ai, arows, bi, brows, 0.0, t2, M, 1) ||
        EXIT_SUCCESS != ozaki_gemm(ctx, stream, tra, trb, M, N, K, 1.0, as, arows, bs, brows, 0.0, t3, M, 1)) {
      return EXIT_FAILURE;
    }
    result = lxs_oz_ccombine(dtype, t1, t2, t3, M, N, alpha, beta, c, ldc, hs);
  }
  return (0 == result ? EXIT_SUCCESS : lxs_error(EXIT_FAILURE, "ozaki_gemm_complex: kernel launch failed"));
}

LIBXSTREAM_API int ozaki_dgemm_native(libxstream_stream_t* stream, char transa, char transb, int M, int N, int K,
  double alpha, const double* a, int lda, const double* b, int ldb, double beta, double* c, int ldc) {
  const int ta = oz_trans(transa), tb = oz_trans(transb);
  hipStream_t hs;
  void* token;
  int result;
  if (0 > ta || 0 > tb || 0 > M || 0 > N || 0 > K) return lxs_error(EXIT_FAILURE, "ozaki_dgemm_native: invalid argument");
  if (lda < (1 > (0 == ta ? M : K) ? 1 : (0 == ta ? M : K)) || ldb < (1 > (0 == tb ? K : N) ? 1 : (0 == tb ? K : N)) ||
      ldc < (1 > M ? 1 : M)) {
    return lxs_error(EXIT_FAILURE, "ozaki_dgemm_native: leading dimension too small");
  }
  if (0 == M || 0 == N) return EXIT_SUCCESS;
  if (0 != lxs_ensure_device()) return EXIT_FAILURE;
  hs = lxs_stream_hip(stream);
  token = lxs_profile_begin(hs);
  result = lxs_dgemm(a, 0 == ta ? 1 : lda, 0 == ta ? lda : 1, b, 0 == tb ? 1 : ldb, 0 == tb ? ldb : 1, M, N, K, alpha,
    beta, c, ldc, hs);
  lxs_profile_end(token, hs, "dgemm", 2.0 * M * N * K, 8.0 * ((double)M * K + (double)K * N + 2.0 * M * N));
  return (0 == result ? EXIT_SUCCESS : lxs_error(EXIT_FAILURE, "ozaki_dgemm_native: kernel launch failed"));
}

LIBXSTREAM_API int ozaki_stack(ozaki_context_t* ctx, libxstream_stream_t* stream, const int* dev_stack, int stack_size,
  int m, int n, int k, const void* dev_a, int na, const void* dev_b, int nb, void* dev_c) {
  const int dtype = (NULL != ctx && 0 != ctx->use_double ? 3 : 1);
  oz_impl_t* impl;
  hipStream_t hs;
  size_t pla, plb, need;
  char* base;
  void* token;
  int s, result;
  if (NULL == ctx || NULL == ctx->impl) return lxs_error(EXIT_FAILURE, "ozaki_stack: context not initialised");
  if (0 > stack_size || 1 > m || 64 < m || 1 > n || 64 < n || 1 > k || 64 < k) {
    return lxs_error(EXIT_FAILURE, "ozaki_stack: invalid argument (blocks of at most 64 x 64 x 64)");
  }
  if (0 == stack_size) return EXIT_SUCCESS;
  if (NULL == dev_stack || NULL == dev_a || NULL == dev_b || NULL == dev_c || 0 >= na || 0 >= nb) {
    return lxs_error(EXIT_FAILURE, "ozaki_stack: invalid argument");
  }
  if (0 != lxs_ensure_device()) return EXIT_FAILURE;
  impl = (oz_impl_t*)ctx->impl;
  hs = lxs_stream_hip(stream);
  s = ctx->nslices;
  pla = oz_up((size_t)na * s * 4096, OZ_ALIGN);
  plb = oz_up((size_t)nb * s * 4096, OZ_ALIGN);
  need = pla + plb + oz_up(sizeof(int) * 64 * (size_t)na, OZ_ALIGN) + oz_up(sizeof(int) * 64 * (size_t)nb, OZ_ALIGN) + OZ_ALIGN;
  if (impl->sarena_size < need) {
    if (NULL != impl->sarena) {
      if (0 != lxs_hip(hipStreamSynchronize(hs), "ozaki_stack: stream synchronize")) return EXIT_FAILURE;
      (void)hipFree(impl->sarena);
      impl->sarena = NULL;
      impl->sarena_size = 0;
    }
    if (0 != lxs_hip(hipMalloc(&impl->sarena, need), "ozaki_stack: scratch allocation")) return EXIT_FAILURE;
    impl->sarena_size = need;
  }
  base = (char*)impl->sarena;
  token = lxs_profile_begin(hs);
  result = lxs_ozs_process(dtype, dev_stack, stack_size, m, n, k, dev_a, na, dev_b, nb, dev_c, s, oz_cutoff(ctx),
    (signed char*)base, (int*)(base + pla + plb), (signed char*)(base + pla),
    (int*)(base + pla + plb + oz_up(sizeof(int) * 64 * (size_t)na, OZ_ALIGN)), (int*)(base + need - OZ_ALIGN), hs);
  lxs_profile_end(token, hs, "ozaki_stack", 2.0 * stack_size * m * n * k,
    (double)stack_size * (0 != ctx->use_double ? 8 : 4) * ((double)m * k + (double)k * n + 2.0 * m * n));
  return (0 == result ? EXIT_SUCCESS : lxs_error(EXIT_FAILURE, "ozaki_stack: kernel launch failed"));
}
