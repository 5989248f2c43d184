// MFMA fragment conventions on gfx950 shared by the GEMM kernels.
//
// v_mfma_{f32_16x16x4f32, f64_16x16x4f64}: lane l supplies A[m = l%16][k = l/16]
// and B[k = l/16][n = l%16] (one element each) and receives four elements of
// the 16x16 output in column n = l%16, rows Acc<T>::row(l/16, i), i = 0..3.
#pragma once
#include <hip/hip_runtime.h>

namespace fricc {

typedef __attribute__((ext_vector_type(4))) double d4;
typedef __attribute__((ext_vector_type(4))) float f4;

template <typename T>
struct Acc;
template <>
struct Acc<double> {
  typedef d4 type;
  static __device__ inline d4 mfma(double a, double b, d4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // D register i of lane group g holds row g + 4 i (measured on gfx950)
  static __device__ inline int row(int g, int i) { return g + 4 * i; }
};
template <>
struct Acc<float> {
  typedef f4 type;
  static __device__ inline f4 mfma(float a, float b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // D register i of lane group g holds row 4 g + i
  static __device__ inline int row(int g, int i) { return 4 * g + i; }
};

}  // namespace fricc
