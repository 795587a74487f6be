// Three-part bf16 splitting of FP32 values for the bf16 matrix cores (device
// helpers shared by the complex64 GEMM kernels).
//
// v = hi + mid + lo + O(2^-27 |v|): each part is the round-to-nearest bf16 of
// what the previous ones leave, and the residuals are exact in FP32.  A real
// product x y is then formed from the six partial products above 2^-24 |x y|
// (hi.hi, hi.mid, mid.hi, hi.lo, mid.mid, lo.hi), accumulated in FP32.
#pragma once

#include <hip/hip_runtime.h>

namespace nc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

// two FP32 values rounded to nearest bf16, packed (a in the low half)
__device__ __forceinline__ unsigned bf16_pair(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2{a, b}), bf16x2));
}
// the packed bf16 values as FP32 (a bf16 is the upper half of an FP32)
__device__ __forceinline__ float bf16_lo(unsigned u) { return __builtin_bit_cast(float, u << 16); }
__device__ __forceinline__ float bf16_hi(unsigned u) { return __builtin_bit_cast(float, u & 0xffff0000u); }

// the three parts of 8 values, 8 bf16 (16 bytes) per part
__device__ __forceinline__ void split3(const float (&v)[8], u32x4& H, u32x4& Mid, u32x4& L) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        float a = v[2 * e], b = v[2 * e + 1];
        const unsigned ph = bf16_pair(a, b);
        a -= bf16_lo(ph);
        b -= bf16_hi(ph);
        const unsigned pm = bf16_pair(a, b);
        a -= bf16_lo(pm);
        b -= bf16_hi(pm);
        H[e] = ph;
        Mid[e] = pm;
        L[e] = bf16_pair(a, b);
    }
}
__device__ __forceinline__ void split3(const float (&v)[8], bf16x8& h, bf16x8& m, bf16x8& l) {
    u32x4 H, Mid, L;
    split3(v, H, Mid, L);
    h = __builtin_bit_cast(bf16x8, H);
    m = __builtin_bit_cast(bf16x8, Mid);
    l = __builtin_bit_cast(bf16x8, L);
}

}  // namespace nc
