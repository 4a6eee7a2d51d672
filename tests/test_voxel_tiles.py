"""Host construction of the voxel tiles of the tiled derivation
(cardiotrust_amd/csrc/tiles.cpp), built with g++ and checked on the CPU.

Every table must be a valid input of derive_tiled_kernel whatever the
neighbour table holds: the tiles partition the voxels, no tile exceeds the
voxel or halo limit, each tile's halo is exactly its voxels and their
neighbours, and every slot names the connection's neighbour. On the
120,120-voxel shell the tiles must also be compact (the halo holds about 4
rows per voxel rather than the 27 of unrelated voxels), which is what the
kernel's traffic depends on.
"""
import os
import subprocess

import numpy as np
import pytest

from cardiotrust_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MAX_VOXELS, MAX_HALO = 39, 152


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("tiles") / "tiles_driver")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "cardiotrust_amd", "csrc"),
                    os.path.join(HERE, "native", "tiles_driver.cpp"),
                    os.path.join(ROOT, "cardiotrust_amd", "csrc", "tiles.cpp"), "-o", exe],
                   check=True)
    return exe


def build(driver, tmp_path, nbr, max_voxels=MAX_VOXELS, max_halo=MAX_HALO, raw=False):
    N = nbr.shape[0]
    src = tmp_path / "nbr.i32"
    np.ascontiguousarray(nbr, dtype=np.int32).tofile(src)
    prefix = str(tmp_path / "tiles")
    subprocess.run([driver, str(src), str(N), str(max_voxels), str(max_halo), prefix] +
                   (["raw"] if raw else []), check=True)
    load = lambda name, dt: np.fromfile(f"{prefix}.{name}", dtype=dt)
    return (load("vox_ptr.i32", np.int32), load("vox.i32", np.int32), load("halo_ptr.i32", np.int32),
            load("halo.i32", np.int32), load("slot.u8", np.uint8))


def check(nbr, tables, max_voxels=MAX_VOXELS, max_halo=MAX_HALO):
    vox_ptr, vox, halo_ptr, halo, slot = tables
    N = nbr.shape[0]
    G = vox_ptr.size - 1
    assert vox_ptr[0] == 0 and vox_ptr[-1] == N and halo_ptr[0] == 0 and halo_ptr[-1] == halo.size
    assert np.array_equal(np.sort(vox), np.arange(N))  # a partition
    assert slot.size == 26 * N
    sizes, halos = np.diff(vox_ptr), np.diff(halo_ptr)
    assert sizes.min() >= 1 and sizes.max() <= max_voxels and halos.max() <= max_halo
    for g in range(G):
        own = vox[vox_ptr[g]:vox_ptr[g + 1]]
        h = halo[halo_ptr[g]:halo_ptr[g + 1]]
        assert np.unique(h).size == h.size
        w = nbr[own].ravel()
        assert set(h.tolist()) == set(own.tolist()) | set(w[w >= 0].tolist())
        s = slot[26 * vox_ptr[g]:26 * vox_ptr[g + 1]]
        assert np.all((s == 255) == (w < 0))
        assert np.array_equal(h[s[w >= 0]], w[w >= 0])
    return G, halo.size / N


def test_tiles_of_the_120k_shell_are_compact(driver, tmp_path):
    g = synth.ellipsoid_shell()
    nbr = g.neighbours.reshape(-1, 26)
    G, per_voxel = check(nbr, build(driver, tmp_path, nbr))
    assert g.num_voxels == 120120
    assert per_voxel < 4.6, per_voxel  # halo rows per voxel
    assert G * MAX_VOXELS < 1.25 * g.num_voxels  # tiles mostly full


@pytest.mark.parametrize("shape,fill,seed", [((10, 10, 10), 1.0, 0), ((9, 4, 6), 0.45, 8),
                                            ((7, 6, 5), 0.6, 5), ((1, 1, 1), 1.0, 0)])
def test_tiles_of_grid_blobs(driver, tmp_path, shape, fill, seed):
    mask = synth.uniform(seed, 99, int(np.prod(shape))).reshape(shape) < fill
    mask.flat[0] = True
    nbr = synth.grid_geometry(mask, 2.5).neighbours.reshape(-1, 26)
    check(nbr, build(driver, tmp_path, nbr))


def test_tiles_of_a_table_that_is_no_grid(driver, tmp_path):
    """Random neighbours (asymmetric, repeated, missing): still a valid
    partition within both limits, and tighter limits split every block."""
    N = 500
    rng = np.random.default_rng(3)
    nbr = rng.integers(-1, N, size=(N, 26)).astype(np.int32)
    check(nbr, build(driver, tmp_path, nbr))
    check(nbr, build(driver, tmp_path, nbr, max_voxels=3, max_halo=60), 3, 60)


def test_dispatch_order_keeps_each_xcd_range_and_puts_large_tiles_first(driver, tmp_path):
    """The library uploads the tiles in dispatch order: launch block b runs
    tile xcd_block(b), so XCD k takes one contiguous range of tile indices
    (the first G mod 8 ranges one tile longer) in increasing order. Within
    each range the tiles must be the builder's tiles of that range, largest
    first (ties in builder order)."""
    g = synth.ellipsoid_shell()
    nbr = g.neighbours.reshape(-1, 26)
    raw = build(driver, tmp_path, nbr, raw=True)
    got = build(driver, tmp_path, nbr)
    check(nbr, got)
    G = raw[0].size - 1
    assert got[0].size - 1 == G
    key = lambda t, i: tuple(t[1][t[0][i]:t[0][i + 1]])  # a tile: its voxels in order
    q, r = divmod(G, 8)
    begin = 0
    for x in range(8):
        end = begin + q + (1 if x < r else 0)
        want = sorted(range(begin, end), key=lambda i: (-(raw[0][i + 1] - raw[0][i]), i))
        assert [key(got, i) for i in range(begin, end)] == [key(raw, i) for i in want], x
        begin = end
    assert begin == G
