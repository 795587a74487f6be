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

// Two-part fp16 splitting of FP32 values already multiplied by a power of two
// that keeps them at most 2^14 in magnitude: v = hi + lo + O(2^-23 |v|) for v
// in the fp16 normal range (hi the round-to-nearest fp16 of v, lo that of the
// exact FP32 residual); below it the parts keep an absolute accuracy of 2^-25.
// A real product x y is then formed from the three partial products hi.hi,
// hi.lo, lo.hi (the dropped lo.lo is below 2^-22 |x y|).
namespace nc {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

// two FP32 values rounded to nearest fp16, packed (a in the low half)
__device__ __forceinline__ unsigned f16_pair(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2{a, b}), f16x2));
}
// the packed fp16 values as FP32
__device__ __forceinline__ f32x2 f16_unpack(unsigned u) {
    return __builtin_convertvector(__builtin_bit_cast(f16x2, u), f32x2);
}
// hi and lo parts of the pair (a, b)
__device__ __forceinline__ void split2_pair(float a, float b, unsigned& h, unsigned& l) {
    h = f16_pair(a, b);
    const f32x2 hv = f16_unpack(h);
    l = f16_pair(a - hv.x, b - hv.y);
}
// the two parts of 8 values, 8 fp16 (16 bytes) per part
__device__ __forceinline__ void split2(const float (&v)[8], u32x4& H, u32x4& L) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        unsigned h, l;
        split2_pair(v[2 * e], v[2 * e + 1], h, l);
        H[e] = h;
        L[e] = l;
    }
}

}  // namespace nc
