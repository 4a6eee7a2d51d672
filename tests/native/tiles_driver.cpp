// Runs build_voxel_tiles on a neighbour table for tests/test_voxel_tiles.py.
// Usage: tiles_driver <nbr.i32> <N> <max_voxels> <max_halo> <out prefix> [raw]
// (with `raw` the tiles keep the builder's order instead of the dispatch order)
// Reads N * 26 int32 neighbours; writes <prefix>.{vox_ptr,vox,halo_ptr,halo}.i32
// and <prefix>.slot.u8.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "tiles.h"

template <typename V>
static bool write(const std::string &path, const std::vector<V> &v) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const size_t ok = v.empty() ? 0 : std::fwrite(v.data(), sizeof(V), v.size(), f);
    std::fclose(f);
    return ok == v.size();
}

int main(int argc, char **argv) {
    if (argc != 6 && argc != 7) return 2;
    const int N = std::atoi(argv[2]);
    std::vector<int32_t> nbr(size_t(N) * 26);
    FILE *f = std::fopen(argv[1], "rb");
    if (!f || std::fread(nbr.data(), sizeof(int32_t), nbr.size(), f) != nbr.size()) return 3;
    std::fclose(f);
    ctrt::VoxelTiles t;
    ctrt::build_voxel_tiles(nbr, N, std::atoi(argv[3]), std::atoi(argv[4]), 0xFF, t);
    if (argc == 6) ctrt::order_tiles_for_dispatch(t, 8);  // as the library uploads them
    const std::string p = argv[5];
    const bool ok = write(p + ".vox_ptr.i32", t.vox_ptr) && write(p + ".vox.i32", t.vox) &&
                    write(p + ".halo_ptr.i32", t.halo_ptr) && write(p + ".halo.i32", t.halo) &&
                    write(p + ".slot.u8", t.slot);
    return ok ? 0 : 4;
}
