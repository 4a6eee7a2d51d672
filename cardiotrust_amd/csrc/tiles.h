// Voxel tiles of the tiled derivation (derive_tiled_kernel in kernels.hip).
#pragma once
#include <cstdint>
#include <vector>

namespace ctrt {

// Spatially compact groups of at most `max_voxels` voxels whose halo (the
// voxels and all their neighbours) has at most `max_halo` entries. Tile g owns
// the tile-order positions [vox_ptr[g], vox_ptr[g+1]) holding the voxels
// vox[.] and the halo voxels halo[halo_ptr[g] .. halo_ptr[g+1]); slot[26 p + d]
// is the halo index of the neighbour of connection d of the voxel at position
// p, or `no_slot` if that neighbour is missing.
struct VoxelTiles {
    std::vector<int32_t> vox_ptr, vox, halo_ptr, halo;
    std::vector<uint8_t> slot;
    int32_t tiles() const { return int32_t(vox_ptr.empty() ? 0 : vox_ptr.size() - 1); }
};

// Builds the tiles from the neighbour table nbr (N, 26), -1 = missing. Voxel
// coordinates are recovered from the connections: connection d points to the
// neighbour at the grid offset of delay index d (allpass.rs:603-618), so a walk
// over the connections places every voxel of a connected component. Voxels are
// grouped by 3 x 3 x 4 blocks of those coordinates, and consecutive blocks are
// merged while both limits hold; any block that breaks one is split. Every
// voxel lands in exactly one tile whatever the table holds; only the
// locality of the tiles depends on it being a grid.
void build_voxel_tiles(const std::vector<int32_t> &nbr, int32_t N, int max_voxels, int max_halo,
                       uint8_t no_slot, VoxelTiles &out);

// Reorders the tiles for dispatch over `groups` equal ranges of tile
// indices (the device's XCDs: launch block b runs tile xcd_block(b), so each
// XCD takes one contiguous range in increasing order): within each range the
// tiles with more voxels come first, so the shortest run last and the tail
// of the launch is short. The ranges keep their tiles, and with them the
// spatial locality the L2 of each XCD shares.
void order_tiles_for_dispatch(VoxelTiles &t, int groups);

}  // namespace ctrt
