// MFMA helpers shared by the probabilistic NFA-64 kernels (gfx950):
// operand types, the 16x16x32 FP16 / BF16 MFMA, the column maximum of an
// MFMA result and the FP32 -> operand split. Header-only, internal linkage.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace pnfadev {

template <bool BF16>
struct Lp;

template <>
struct Lp<false> {
    typedef f16x8 vec;
    typedef _Float16 scalar;
    __device__ __forceinline__ static f32x4 mfma(vec a, vec b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Lp<true> {
    typedef bf16x8 vec;
    typedef __bf16 scalar;
    __device__ __forceinline__ static f32x4 mfma(vec a, vec b, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};

// Largest value over the four lanes l, l^16, l^32, l^48 (the lanes that hold
// one column of an MFMA result), with the gfx950 lane-swap instructions.
__device__ __forceinline__ float column_max(float v) {
    unsigned u = __float_as_uint(v);
    const auto a = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    v = fmaxf(__uint_as_float(a[0]), __uint_as_float(a[1]));
    u = __float_as_uint(v);
    const auto b = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    return fmaxf(__uint_as_float(b[0]), __uint_as_float(b[1]));
}

// Sum over the same four lanes.
__device__ __forceinline__ float column_sum(float v) {
    unsigned u = __float_as_uint(v);
    const auto a = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    v = __uint_as_float(a[0]) + __uint_as_float(a[1]);
    u = __float_as_uint(v);
    const auto b = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}

// Eight FP32 values -> MFMA operand: the high half (FP32 truncated to the
// operand's precision, converted exactly) and, when split, the rounded rest.
// FP16 high halves come two at a time from the round-toward-zero pack, which
// truncates in the FP16 subnormal range too (a mantissa mask does not), so
// the rest is always x minus the exact high half.
template <bool BF16, bool SPLIT>
__device__ __forceinline__ void split8(const float (&x)[8], typename Lp<BF16>::vec& hi, typename Lp<BF16>::vec& lo) {
    typedef typename Lp<BF16>::scalar scalar;
    if constexpr (!BF16 && SPLIT) {
        typedef __fp16 h2 __attribute__((ext_vector_type(2)));
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        u32x4 hw, lw;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const h2 t = __builtin_amdgcn_cvt_pkrtz(x[2 * p], x[2 * p + 1]);
            const uint32_t bits = __builtin_bit_cast(uint32_t, t);
            // rest = round(x - high) into each half by one mixed-precision FMA
            // (FP16 source, FP32 sum, one rounding to FP16): the subtraction is
            // exact, so this is the same value as converting, subtracting and
            // rounding, in one instruction per element instead of three
            uint32_t rest;
            asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(rest) : "v"(bits), "v"(x[2 * p]));
            asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
                : "+v"(rest) : "v"(bits), "v"(x[2 * p + 1]));
            hw[p] = bits;
            lw[p] = rest;
        }
        hi = __builtin_bit_cast(typename Lp<BF16>::vec, hw);
        lo = __builtin_bit_cast(typename Lp<BF16>::vec, lw);
    } else {
        constexpr unsigned kMask = BF16 ? 0xffff0000u : 0xffffe000u;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            if constexpr (SPLIT) {
                const float h = __uint_as_float(__float_as_uint(x[e]) & kMask);
                hi[e] = (scalar)h;
                lo[e] = (scalar)(x[e] - h);
            } else {
                hi[e] = (scalar)x[e];
            }
        }
    }
}

}  // namespace pnfadev
