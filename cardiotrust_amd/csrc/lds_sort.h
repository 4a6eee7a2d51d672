// Workgroup sort of 64-bit keys in LDS (set-up of the resident kernels).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ct {

// Connections of one slab are sorted once per launch: kSortSize (a power of
// two >= kResMaxVoxels * 26) 8-byte keys over the LDS that later holds the
// slab's filter values. Sorted slabs keep each connection's delay in a byte
// table, so a slab sorts only when all its delays are <= kSortedMaxDelay.
constexpr int kSortSize = 16384;
constexpr int kSortedMaxDelay = 255;

// Ascending bitonic sort of kSize keys by all kThreads threads of the
// workgroup (every thread must call it).
template <int kThreads, int kSize>
__device__ __forceinline__ void sort_keys_lds(uint64_t *keys) {
    static_assert((kSize & (kSize - 1)) == 0, "power of two");
    for (int k = 2; k <= kSize; k <<= 1) {
        for (int jj = k >> 1; jj > 0; jj >>= 1) {
            for (int i = threadIdx.x; i < kSize / 2; i += kThreads) {
                const int lo = 2 * i - (i & (jj - 1));  // i-th index with bit jj clear
                const int hi = lo + jj;
                const uint64_t a = keys[lo], c = keys[hi];
                if ((a > c) == ((lo & k) == 0)) {
                    keys[lo] = c;
                    keys[hi] = a;
                }
            }
            __syncthreads();
        }
    }
}

}  // namespace ct
