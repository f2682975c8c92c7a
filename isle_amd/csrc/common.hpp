// Shared definitions for the gfx950 fermion-action kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace isle {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

template <typename T> struct Acc4;
template <> struct Acc4<double> { typedef dbl4 type; };
template <> struct Acc4<float> { typedef flt4 type; };

// One 16x16x4 MFMA step. Lane l supplies A[l%16][l/16] and B[l/16][l%16].
// Accumulator entry r (0..3) of lane l is C[acc_row<T>(l/16, r)][l%16]:
//   f32: row 4*(l/16) + r      f64: row (l/16) + 4*r
// (measured on gfx950 with an in-bounds probe of both builtins).
__device__ __forceinline__ dbl4 mfma16(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ flt4 mfma16(float a, float b, flt4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

template <typename T> __device__ __forceinline__ int acc_row(int lk, int r);
template <> __device__ __forceinline__ int acc_row<float>(int lk, int r) { return 4 * lk + r; }
template <> __device__ __forceinline__ int acc_row<double>(int lk, int r) { return lk + 4 * r; }

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        throw HipError(std::string(what) + ": " + hipGetErrorString(e));
    }
}

#define ISLE_HIP(call) ::isle::hip_check((call), #call)

inline unsigned cdiv(long a, long b) { return static_cast<unsigned>((a + b - 1) / b); }

}  // namespace isle
