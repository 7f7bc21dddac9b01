// ksh_mfma.hpp -- 16x16x4 MFMA tiles for gfx950 in FP64 and FP32.
//
// Operands (both precisions): lane l supplies A[l%16][l/16] and B[l/16][l%16] of D = A B + C.
// Result: lane l holds D[row(l, v)][l%16], v = 0..3, where the row map differs by precision
// (measured on gfx950 with an in-bounds layout probe):
//   v_mfma_f64_16x16x4f64 : row = l/16 + 4 v
//   v_mfma_f32_16x16x4f32 : row = 4 (l/16) + v
#pragma once

#include <hip/hip_runtime.h>

namespace ksh {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

template <typename T> struct Mfma16;

template <> struct Mfma16<double> {
    typedef v4d acc_t;
    __device__ static acc_t run(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int lane, int v) { return lane / 16 + 4 * v; }
};

template <> struct Mfma16<float> {
    typedef v4f acc_t;
    __device__ static acc_t run(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int lane, int v) { return 4 * (lane / 16) + v; }
};

}  // namespace ksh
