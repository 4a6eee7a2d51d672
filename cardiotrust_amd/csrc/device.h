// Small device helpers shared by the kernel files.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ct {

// Rust f32::signum: NaN stays NaN, -0 -> -1, +0 -> +1.
__device__ __forceinline__ float signum(float v) { return isnan(v) ? v : copysignf(1.0f, v); }

// powi(x, 5) as LLVM expands it for a constant exponent: x * (x^2)^2.
__device__ __forceinline__ float powi5(float v) {
    const float sq = v * v;
    const float q = sq * sq;
    return v * q;
}

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2). Remap the launch order so that XCD k runs one contiguous range of
// blocks: neighbouring voxels, which read the same state rows, then share an
// L2. A bijection on [0, nb).
constexpr int kNumXcds = 8;
__device__ __forceinline__ int32_t xcd_block(int32_t b, int32_t nb) {
    const int32_t xcd = b % kNumXcds, local = b / kNumXcds;
    const int32_t q = nb / kNumXcds, r = nb % kNumXcds;
    return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

// Raw buffer resources (byte offsets, stride 0): word 3 of the descriptor
// with the 32-bit data format, and the cache-policy operand that makes a
// buffer access coherent at agent scope (sc1: it does not hit a stale copy in
// this XCD's L2).
constexpr int32_t kBufferWord3 = 0x00020000;
constexpr int32_t kCacheSC1 = 16;

}  // namespace ct
