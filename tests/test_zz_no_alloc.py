"""ct_results_calculate_final_metrics_noalloc against
ct_results_calculate_final_metrics, through the C ABI.

The two entry points run the same kernels; the first stages its inputs and
keeps its maxima and histograms in the library's fixed device scratch instead
of an allocation. The counts behind the four metrics are integers and the
ratios one f32 division each, so the bar is bit-exact equality. Two calls in
a row, the second with a smaller problem (fewer voxels, fewer cells), check
that nothing from the first call's scratch leaks into the second; a third
call past the scratch's size must fail cleanly before any device work.
"""
import sys

import numpy as np
import pytest

from cardiotrust_amd import synth

pytestmark = pytest.mark.gpu

PATHOLOGICAL, VENTRICLE = 6, 5


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    from cardiotrust_amd.gpu import GPU
    g = GPU(0)
    yield g
    g.close()


def ground_truth(geometry, seed):
    """VoxelType ordinals per cell (a seeded Pathological sphere, Ventricle
    elsewhere, None outside the mask, two Pathological cells without a
    voxel) and the voxel numbers (3 per voxel, -1 for None)."""
    types = np.zeros(geometry.mask.size, np.uint32)
    numbers = np.where(geometry.numbers.ravel() >= 0, 3 * geometry.numbers.ravel(), -1)
    cells = np.flatnonzero(geometry.mask.ravel())
    region = synth.sphere_region(geometry, seed, radius_fraction=0.3)
    types[cells] = np.where(region, PATHOLOGICAL, VENTRICLE)
    types[np.flatnonzero(~geometry.mask.ravel())[:2]] = PATHOLOGICAL
    return types, numbers.astype(np.int32)


def simulated(gpu, geometry, seed, T):
    from cardiotrust_amd.gpu import ResultsGPU
    p = synth.make_problem(geometry, num_steps=T, max_delay=12, seed=seed, gain_sigma=0.3).truth()
    # drive every state for a few steps so that the maxima spread over the thresholds
    p.control_function = synth.uniform(seed, 40, T).astype(np.float32)
    p.control_matrix = (0.4 * synth.uniform(seed, 41, p.num_states)).astype(np.float32)
    r = ResultsGPU(gpu, p)
    r.simulate()
    return r


def assert_bits(got, want):
    for k in want:
        assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), k


def test_noalloc_final_metrics_equal_the_allocating_entry_point(gpu):
    from cardiotrust_amd.gpu import CTError
    cases = [("blob 9000 voxels", synth.blob(31, 9000), 31, 12),
             ("blob 700 voxels", synth.blob(32, 700), 32, 10)]
    assert cases[1][1].num_voxels < cases[0][1].num_voxels
    assert cases[1][1].mask.size < cases[0][1].mask.size
    for name, geom, seed, T in cases:
        note(f"case: {name} ({geom.num_voxels} voxels, {geom.mask.size} cells), T={T}")
        r = simulated(gpu, geom, seed, T)
        types, numbers = ground_truth(geom, seed)
        want = r.calculate_final(types, numbers)
        got = r.calculate_final(types, numbers, no_alloc=True)
        r.close()
        assert_bits(got, want)
        assert len(set(want["dice_score_over_threshold"].tolist())) > 10  # not a flat sweep
    note("case: cells beyond the scratch size (host check, no device work)")
    r = simulated(gpu, synth.blob(33, 300), 33, 10)
    cells = 1_100_000  # 8.8 MB of staged cell arrays
    with pytest.raises(CTError, match="scratch"):
        r.calculate_final(np.zeros(cells, np.uint32), np.full(cells, -1, np.int32),
                          no_alloc=True)
    r.close()
