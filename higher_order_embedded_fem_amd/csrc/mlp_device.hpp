// Device helpers of the learned coarse correction: the per-aggregate network
// 64 -> 128 -> 128 -> 64 with GELU, BF16 weights and activations as products' inputs, FP32
// accumulation and FP32 biases.
//
// Parameter layout (FP32 master copy and BF16 weight copy share the offsets):
//   W1 [128][64], b1 [128], W2 [128][128], b2 [128], W3 [64][128], b3 [64]
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hoefem {

constexpr int LC_IN = 64, LC_HID = 128, LC_OUT = 64;
constexpr int LC_W1 = 0;
constexpr int LC_B1 = LC_W1 + LC_HID * LC_IN;
constexpr int LC_W2 = LC_B1 + LC_HID;
constexpr int LC_B2 = LC_W2 + LC_HID * LC_HID;
constexpr int LC_W3 = LC_B2 + LC_HID;
constexpr int LC_B3 = LC_W3 + LC_OUT * LC_HID;
constexpr int LC_PARAMS = LC_B3 + LC_OUT;  // 33088

// round-to-nearest-even FP32 -> BF16 (finite inputs) and back
__host__ __device__ inline uint16_t bf16_bits(float x) {
    union {
        float f;
        uint32_t u;
    } c{x};
    const uint32_t u = c.u;
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__host__ __device__ inline float bf16_float(uint16_t b) {
    union {
        uint32_t u;
        float f;
    } c{(uint32_t)b << 16};
    return c.f;
}
__host__ __device__ inline float bf16_round(float x) { return bf16_float(bf16_bits(x)); }

// GELU(x) = x Phi(x) with the normal CDF Phi (erf form), and its derivative
__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

}  // namespace hoefem
