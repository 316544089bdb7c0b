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
