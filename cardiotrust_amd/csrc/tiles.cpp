// Voxel tiles of the tiled derivation: host-side construction.
#include "tiles.h"

#include <algorithm>
#include <array>

namespace ctrt {

namespace {

// Grid offset (dx, dy, dz) of delay index d: offset index o = d, skipping the
// voxel itself at o = 13, with o = 9 (dx + 1) + 3 (dy + 1) + (dz + 1)
// (allpass.rs:582-618).
std::array<int32_t, 3> offset_of(int d) {
    const int o = d < 13 ? d : d + 1;
    return {o / 9 - 1, (o / 3) % 3 - 1, o % 3 - 1};
}

int32_t floor_div(int32_t a, int32_t b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

}  // namespace

void build_voxel_tiles(const std::vector<int32_t> &nbr, int32_t N, int max_voxels, int max_halo,
                       uint8_t no_slot, VoxelTiles &out) {
    out = VoxelTiles{};
    if (N <= 0) return;
    std::array<std::array<int32_t, 3>, 26> off;
    for (int d = 0; d < 26; ++d) off[d] = offset_of(d);

    // coordinates by a walk over the connections, one origin per component
    std::vector<std::array<int32_t, 4>> key(N);  // component, x, y, z
    std::vector<uint8_t> seen(N, 0);
    std::vector<int32_t> queue;
    queue.reserve(N);
    int32_t comp = 0;
    for (int32_t s = 0; s < N; ++s) {
        if (seen[s]) continue;
        seen[s] = 1;
        key[s] = {comp, 0, 0, 0};
        queue.clear();
        queue.push_back(s);
        for (size_t qi = 0; qi < queue.size(); ++qi) {
            const int32_t u = queue[qi];
            for (int d = 0; d < 26; ++d) {
                const int32_t w = nbr[size_t(u) * 26 + d];
                if (w < 0 || w >= N || seen[w]) continue;
                seen[w] = 1;
                key[w] = {comp, key[u][1] + off[d][0], key[u][2] + off[d][1], key[u][3] + off[d][2]};
                queue.push_back(w);
            }
        }
        ++comp;
    }
    // blocks of 3 x 3 x 4 in (x, y, z); voxels of a block in index order
    for (int32_t v = 0; v < N; ++v) {
        key[v][1] = floor_div(key[v][1], 3);
        key[v][2] = floor_div(key[v][2], 3);
        key[v][3] = floor_div(key[v][3], 4);
    }
    std::vector<int32_t> order(N);
    for (int32_t v = 0; v < N; ++v) order[v] = v;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t a, int32_t b) { return key[a] < key[b]; });

    // merge blocks into tiles under both limits
    std::vector<int32_t> mark(N, -1), fresh(N, -1), slot_of(N, 0);
    int32_t tile = 0, nvox = 0, nhalo = 0, stamp = 0;
    out.vox.reserve(N);
    out.slot.reserve(size_t(N) * 26);
    out.vox_ptr.push_back(0);
    out.halo_ptr.push_back(0);
    const auto halo_of = [&](int32_t u, auto &&visit) {
        visit(u);
        for (int d = 0; d < 26; ++d) {
            const int32_t w = nbr[size_t(u) * 26 + d];
            if (w >= 0 && w < N) visit(w);
        }
    };
    // halo entries that voxels[b, e) add to the open tile
    const auto added = [&](size_t b, size_t e) {
        ++stamp;
        int32_t count = 0;
        for (size_t i = b; i < e; ++i)
            halo_of(order[i], [&](int32_t x) {
                if (mark[x] != tile && fresh[x] != stamp) {
                    fresh[x] = stamp;
                    ++count;
                }
            });
        return count;
    };
    const auto close = [&]() {
        if (nvox == 0) return;
        out.vox_ptr.push_back(int32_t(out.vox.size()));
        out.halo_ptr.push_back(int32_t(out.halo.size()));
        ++tile;
        nvox = nhalo = 0;
    };
    const auto add = [&](int32_t u) {
        halo_of(u, [&](int32_t x) {
            if (mark[x] != tile) {
                mark[x] = tile;
                slot_of[x] = nhalo++;
                out.halo.push_back(x);
            }
        });
        out.vox.push_back(u);
        for (int d = 0; d < 26; ++d) {
            const int32_t w = nbr[size_t(u) * 26 + d];
            out.slot.push_back(w >= 0 && w < N ? uint8_t(slot_of[w]) : no_slot);
        }
        ++nvox;
    };
    for (size_t b = 0; b < size_t(N);) {
        size_t e = b + 1;
        while (e < size_t(N) && key[order[e]] == key[order[b]]) ++e;
        const int32_t size = int32_t(e - b);
        bool whole = size <= max_voxels;
        if (whole && (nvox + size > max_voxels || nhalo + added(b, e) > max_halo)) {
            close();
            whole = added(b, e) <= max_halo;
        }
        if (whole) {
            for (size_t i = b; i < e; ++i) add(order[i]);
        } else {
            for (size_t i = b; i < e; ++i) {
                if (nvox + 1 > max_voxels || nhalo + added(i, i + 1) > max_halo) close();
                add(order[i]);
            }
        }
        b = e;
    }
    close();
}

void order_tiles_for_dispatch(VoxelTiles &t, int groups) {
    const int32_t G = t.tiles();
    if (G <= 1 || groups < 1) return;
    std::vector<int32_t> perm(G);
    for (int32_t g = 0; g < G; ++g) perm[g] = g;
    const int32_t q = G / groups, r = G % groups;
    for (int32_t x = 0, begin = 0; x < groups; ++x) {
        const int32_t len = q + (x < r ? 1 : 0);
        std::stable_sort(perm.begin() + begin, perm.begin() + begin + len, [&](int32_t a, int32_t b) {
            return t.vox_ptr[a + 1] - t.vox_ptr[a] > t.vox_ptr[b + 1] - t.vox_ptr[b];
        });
        begin += len;
    }
    VoxelTiles o;
    o.vox_ptr.push_back(0);
    o.halo_ptr.push_back(0);
    for (int32_t g : perm) {
        o.vox.insert(o.vox.end(), t.vox.begin() + t.vox_ptr[g], t.vox.begin() + t.vox_ptr[g + 1]);
        o.slot.insert(o.slot.end(), t.slot.begin() + size_t(t.vox_ptr[g]) * 26,
                      t.slot.begin() + size_t(t.vox_ptr[g + 1]) * 26);
        o.halo.insert(o.halo.end(), t.halo.begin() + t.halo_ptr[g],
                      t.halo.begin() + t.halo_ptr[g + 1]);
        o.vox_ptr.push_back(int32_t(o.vox.size()));
        o.halo_ptr.push_back(int32_t(o.halo.size()));
    }
    t = std::move(o);
}

}  // namespace ctrt
