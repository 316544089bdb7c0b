/*
 * Matrix-core helpers shared by the stack kernels (HIP device code, gfx950).
 *
 * MFMA 16x16x4 with one operand element per lane: lane l holds A[l % 16][l / 16] of a 16x4 A
 * fragment and B[l / 16][l % 16] of a 4x16 B fragment. Output tile 16x16, lane l, register r,
 * column l % 16: FP64 row = l/16 + 4r, FP32 row = 4(l/16) + r (checked on gfx950 with a probe).
 */
#ifndef LXS_MFMA_H
#define LXS_MFMA_H

#include <hip/hip_runtime.h>

namespace lxs {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename T>
struct mfma;
template <>
struct mfma<double> {
  typedef d4_t acc_t;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct mfma<float> {
  typedef f4_t acc_t;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

/* 4x4-block MFMA (checked on gfx950 with probes). FP64 v_mfma_f64_4x4x4_4b: four blocks of 4x4x4; lane l feeds
   k = l / 16 of block (l / 4) % 4, A row l % 4 and B column l % 4, and holds the block's result for A row l / 16
   and B column l % 4 (one register). FP32 v_mfma_f32_4x4x1_16b: sixteen blocks of 4x4x1; lane l feeds block
   l / 4 (its one k), A row l % 4 and B column l % 4, and holds the block's column l % 4 in its four registers
   (register r: row r). klane: the k within 16 that lane l feeds when the blocks of one instruction take
   consecutive k of the same C block. */
template <typename T>
struct mfma4;
template <>
struct mfma4<double> {
  typedef double acc_t;
  static constexpr bool FP32 = false;
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int klane(int lane) { return 4 * ((lane >> 2) & 3) + (lane >> 4); }
};
template <>
struct mfma4<float> {
  typedef f4_t acc_t;
  static constexpr bool FP32 = true;
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int klane(int lane) { return lane >> 2; }
};

/* zero operands for fragment lanes past k (read-only) */
__device__ const double zeros_f64[2] = {0.0, 0.0};
__device__ const float zeros_f32[4] = {0.0f, 0.0f, 0.0f, 0.0f};
template <typename T>
__device__ __forceinline__ const T* zeros();
template <>
__device__ __forceinline__ const double* zeros<double>() {
  return zeros_f64;
}
template <>
__device__ __forceinline__ const float* zeros<float>() {
  return zeros_f32;
}

/* load through the global address space: fragment pointers may point at the operand blocks or at
   the zero array, and a generic (flat) load would force the compiler to wait for all loads */
template <typename T>
__device__ __forceinline__ T gload(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

/* two consecutive elements at the element type's alignment (one 16-byte load for FP64, 8-byte for FP32;
   the operand blocks only guarantee the element alignment) */
typedef double d2u_t __attribute__((ext_vector_type(2), aligned(8)));
typedef float f2u_t __attribute__((ext_vector_type(2), aligned(4)));
template <typename T>
struct pair_of;
template <>
struct pair_of<double> {
  typedef d2u_t type;
};
template <>
struct pair_of<float> {
  typedef f2u_t type;
};
template <typename T>
__device__ __forceinline__ typename pair_of<T>::type gload2(const T* p) {
  return *(const __attribute__((address_space(1))) typename pair_of<T>::type*)p;
}

/* Paired tiles: output tiles 2p and 2p+1 of one dimension interleave the 32 rows (or columns) from 32p on,
   so that one two-element load per lane feeds both tiles. Lane l of the pair loads the elements at
   s = min(32p + 2l, extent - 2) and s + 1 (clamped so both stay inside the extent) for tiles 2p and 2p+1.
   Output row l of tile 2p+q is then element s + q; it is written by exactly one lane: an unclamped pair,
   or for an odd extent the lane whose pair would start at extent - 1 (tile 2p+1 holds that last element). */
__device__ __forceinline__ int pair_start(int p, int l, int extent) { return min(32 * p + 2 * l, extent - 2); }
__device__ __forceinline__ bool pair_owner(int p, int q, int l, int extent) {
  const int base = 32 * p + 2 * l;
  return base <= extent - 2 || (1 == q && base == extent - 1);
}

/* Long C-runs are cut into pieces at the segment boundaries of the stack (multiples of the
   segment size bs, where one work-group's entries begin). Boundary entry i starts a new piece
   when the entry tsplit places earlier has the same C block ci: in a stack sorted by C, the run
   through i has then gone on for more than tsplit entries. The two owners that meet at i both
   evaluate this function of the stack alone, so each piece has exactly one owner; pieces of a cut
   run update their C block with atomic adds. tsplit <= 0: runs are never cut. */
__device__ __forceinline__ bool lxs_cut(const int* stack, int ps, int i, int ci, int tsplit) {
  return 0 < tsplit && tsplit <= i && stack[ps * (i - tsplit) + 2] == ci;
}

} /* namespace lxs */

#endif /*LXS_MFMA_H*/
