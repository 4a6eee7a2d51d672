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

}  // namespace ct
