// Device side of the voxel-tile kernels (derive_tiled_kernel in kernels.hip,
// adjoint_gradient_tiled_kernel in adjoint.hip): the lane of each connection
// and the LDS ring of the tile's halo state rows.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"

namespace ct {

// LDS layout of a tile kernel, in floats: the ring [halo][3][kTileRowStride],
// then one zero row triple [3][kTileRowStride] (the inputs of missing
// connections), then the kernel's per-voxel staging arrays [voxel][3][chunk].
constexpr int kTileRingFloats = kTileMaxHalo * 3 * kTileRowStride;
constexpr int kTileZeroFloats = 3 * kTileRowStride;
constexpr int kTileStageFloats = kTileMaxVoxels * 3 * kTileChunk;
constexpr int kTileQuads = kTileMaxHalo * 3 * (kTileChunk / 4);  // 16-byte loads per chunk
constexpr int kTileQuadsPerThread = (kTileQuads + kTileThreads - 1) / kTileThreads;
static_assert(kTileMaxVoxels * kTileChunk <= kTileThreads, "one staging item per thread");
static_assert(kTileRowStride % 4 == 0, "ring rows take 16-byte stores");
static_assert(kTileMaxHalo < kNoSlot, "halo indices fit the slot bytes");
static_assert(kTileRing % kTileChunk == 0 && (kTileRing & (kTileRing - 1)) == 0,
              "chunks fill ring positions without wrapping");

typedef float float4_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef float float4_a16 __attribute__((ext_vector_type(4), aligned(16)));

// One connection of the tile per lane, in tile order: lane 26 lv + d takes
// direction d of the tile's voxel lv, so the 26 lanes of a voxel share its
// staged values (near-broadcast LDS reads). (Lanes sorted by delay read the
// ring without bank conflicts but scatter those reads; the tiled derivation
// measured 11.1 ms against 9.9 in tile order.) `far`: the delay exceeds
// `max_delay`, and the lane reads its inputs from memory.
struct TileLane {
    int lv, d;          // voxel of the tile, direction
    int32_t v, w, D;    // voxel, neighbour (-1 missing), delay
    int64_t item;       // connection v * 26 + d
    float a;            // coefficient
    int slot;           // halo index of the neighbour in the ring, -1 if not read from it
    bool active, has, far;
};

__device__ __forceinline__ TileLane tile_lane(const DeriveTiles &tl, const int32_t *nbr,
                                              const float *coefs, const int32_t *delays,
                                              int32_t pbeg, int32_t nvox, int32_t max_delay) {
    TileLane L;
    L.lv = threadIdx.x / 26;
    L.d = threadIdx.x - 26 * L.lv;
    L.active = L.lv < nvox;
    L.v = L.active ? tl.vox[pbeg + L.lv] : 0;
    L.item = int64_t(L.v) * 26 + L.d;
    L.w = L.active ? nbr[L.item] : -1;
    L.has = L.w >= 0;
    L.a = L.active ? coefs[L.item] : 0.0f;
    L.D = L.active ? delays[L.item] : 0;
    L.far = L.has && !(L.D >= 0 && L.D <= max_delay);
    L.slot = (L.has && !L.far) ? int(tl.slot[int64_t(pbeg + L.lv) * 26 + L.d]) : -1;
    return L;
}

// The copies of one chunk of every halo row into the ring: this thread's
// 16-byte quads, loaded into registers by fetch() (while the previous chunk
// runs) and written to their ring positions by stage().
struct TileRingCopy {
    int64_t src[kTileQuadsPerThread];  // element of xT at step 0, or -1
    int32_t dst[kTileQuadsPerThread];  // ring offset at position 0 (row r: r * kTileRowStride + q4)
    float4_a4 quad[kTileQuadsPerThread];

    __device__ __forceinline__ void init(const DeriveTiles &tl, int32_t hbeg, int32_t nhalo,
                                         int32_t T) {
#pragma unroll
        for (int k = 0; k < kTileQuadsPerThread; ++k) {
            const int qi = threadIdx.x + k * kTileThreads;
            src[k] = -1;
            dst[k] = 0;
            if (qi < nhalo * 3 * (kTileChunk / 4)) {
                const int r = qi / (kTileChunk / 4), q4 = 4 * (qi % (kTileChunk / 4));
                const int h = r / 3, o = r - 3 * h;
                src[k] = (3LL * tl.halo[hbeg + h] + o) * T + q4;
                dst[k] = r * kTileRowStride + q4;
            }
        }
    }
    // steps [t0, t0 + kTileChunk) of the rows; steps >= T read +0
    __device__ __forceinline__ void fetch(const float *xT, int32_t t0, int32_t T) {
        const bool whole = t0 + kTileChunk <= T;
#pragma unroll
        for (int k = 0; k < kTileQuadsPerThread; ++k) {
            if (src[k] < 0) continue;
            if (whole) {
                quad[k] = *reinterpret_cast<const float4_a4 *>(xT + src[k] + t0);
            } else {
                const int32_t tq = t0 + dst[k] % kTileRowStride;  // the quad's first step
#pragma unroll
                for (int e = 0; e < 4; ++e) quad[k][e] = tq + e < T ? xT[src[k] + t0 + e] : 0.0f;
            }
        }
    }
    // ring position t0 mod kTileRing, and behind the ring for position 0
    __device__ __forceinline__ void stage(float *ring, int32_t t0) const {
        const int pos = t0 & (kTileRing - 1);
#pragma unroll
        for (int k = 0; k < kTileQuadsPerThread; ++k) {
            if (src[k] < 0) continue;
            *reinterpret_cast<float4_a16 *>(ring + dst[k] + pos) = quad[k];
            if (pos == 0) *reinterpret_cast<float4_a16 *>(ring + dst[k] + kTileRing) = quad[k];
        }
    }
};

}  // namespace ct
