// qlx_mfma_core.hpp — operand forms shared by the matrix-core kernels (k_dense_mfma in
// qlx_mfma.hip, k_dsweep in qlx_dsweep.hip).
//
// A dense k-qubit gate U (k = 4, 5) is applied as the real GEMM D = A X (A the real embedding
// of U, X the real and imaginary parts of 32 groups of amplitudes). Three arithmetic forms:
//   * BF16: v_mfma_f32_32x32x16_bf16. A = a_hi + a_mid + a_lo (three BF16 terms, 24 bits),
//     X = x_hi + x_lo (two terms, 16 bits); four products accumulated in FP32.
//   * F16:  v_mfma_f32_32x32x16_f16. A = a_hi + a_lo, X = x_hi + x_lo (two FP16 terms each,
//     22 bits); three products (a_lo x_hi, a_hi x_lo, a_hi x_hi; the dropped a_lo x_lo is
//     <= 2^-22 relative). The amplitudes are scaled by 2^14 before the split: a normalised
//     state has |amplitude| <= 1, so scaled values stay 4x below the FP16 maximum (65504), and
//     the typical amplitude of an n-qubit state (2^-n/2) lies in the FP16 normal range up to
//     n = 56. The scale is a power of two: applying and removing it is exact.
//   * F32:  v_mfma_f32_32x32x2_f32 on FP32 operands.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace qlx {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

enum MfmaForm : int { FORM_BF16 = 0, FORM_F32 = 1, FORM_F16 = 2 };

constexpr float kF16Scale = 16384.0f;
constexpr float kF16Unscale = 1.0f / 16384.0f;

// Four amplitudes (the eight values one lane feeds one 16-bit K-step: real and imaginary
// parts interleaved) split into hi and lo operand words, two values at a time: packed
// conversion (round to nearest), the exact value of the hi part back in FP32, the residual
// (exact: hi and the value share their leading bits), packed conversion of the residual.
__device__ __forceinline__ void split_bf16(const float2* x, bf16x8& hi, bf16x8& lo) {
    u32x4 wh, wl;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const f32x2 v = {x[p].x, x[p].y};
        const uint32_t hb = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
        const f32x2 hv = {__builtin_bit_cast(float, hb << 16), __builtin_bit_cast(float, hb & 0xffff0000u)};
        wh[p] = hb;
        wl[p] = __builtin_bit_cast(uint32_t, __builtin_convertvector(v - hv, bf16x2));
    }
    hi = __builtin_bit_cast(bf16x8, wh);
    lo = __builtin_bit_cast(bf16x8, wl);
}

// The same in FP16, each value multiplied by `scale` first (kF16Scale for values read
// unscaled, 1 for values already scaled).
__device__ __forceinline__ void split_f16(const float2* x, float scale, f16x8& hi, f16x8& lo) {
    u32x4 wh, wl;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const f32x2 v = {x[p].x * scale, x[p].y * scale};
        const f16x2 h = __builtin_convertvector(v, f16x2);
        const f32x2 hv = __builtin_convertvector(h, f32x2);
        wh[p] = __builtin_bit_cast(uint32_t, h);
        wl[p] = __builtin_bit_cast(uint32_t, __builtin_convertvector(v - hv, f16x2));
    }
    hi = __builtin_bit_cast(f16x8, wh);
    lo = __builtin_bit_cast(f16x8, wl);
}

}  // namespace qlx
