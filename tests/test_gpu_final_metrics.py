"""ct_results_calculate_final_metrics against the oracle's calculate_final.

The counts behind the four metrics are integers and the ratios are one f32
division each, so the bar is bit-exact equality. The states are the device's
own (simulated there, then read back for the oracle). Every state is driven
directly (a random control matrix) so the maxima spread over the thresholds;
NaN gains at the last voxel make its states NaN at every step and, over ten
steps, its neighbours' NaN from their delay on, which exercises
max_skipnan's two NaN cases while the far side of the blob stays finite.
"""
import sys
import time

import numpy as np
import pytest

from cardiotrust_amd import synth
from oracle import oracle

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


def ground_truth(geometry, seed, extra_none_cells=3):
    """VoxelType ordinals per cell (a seeded Pathological sphere, Ventricle
    elsewhere, None outside the mask, plus a few Pathological cells without a
    voxel) and the reference's voxel numbers (3 per voxel, -1 for None)."""
    types = np.zeros(geometry.mask.size, np.uint32)
    numbers = np.where(geometry.numbers.ravel() >= 0, 3 * geometry.numbers.ravel(), -1)
    cells = np.flatnonzero(geometry.mask.ravel())
    region = synth.sphere_region(geometry, seed, radius_fraction=0.3)
    types[cells] = np.where(region, PATHOLOGICAL, VENTRICLE)
    empty = np.flatnonzero(~geometry.mask.ravel())[:extra_none_cells]
    types[empty] = PATHOLOGICAL
    return types, numbers.astype(np.int32)


def simulated(gpu, problem):
    from cardiotrust_amd.gpu import ResultsGPU
    r = ResultsGPU(gpu, problem)
    r.simulate()
    return r, r.update_from_gpu(["system_states"])["system_states"]


def assert_bits(got, want):
    for k in oracle.FINAL_METRICS:
        assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), \
            (k, np.flatnonzero(got[k] != want[k])[:5])


@pytest.mark.parametrize("seed", [5, 6])
def test_final_metrics_match_the_oracle_with_nan_states(gpu, seed):
    note(f"case: blob of 1500 voxels, 10 steps, NaN gains at the last voxel, seed {seed}")
    T = 10
    g = synth.blob(300 + seed, 1500)
    p = synth.make_problem(g, num_steps=T, max_delay=12, seed=300 + seed, gain_sigma=0.3).truth()
    N = p.num_voxels
    p.gains = p.gains.copy()
    p.gains[3 * (N - 1):, :] = np.nan
    p.control_function = synth.uniform(seed, 40, T).astype(np.float32)
    p.control_matrix = (0.4 * synth.uniform(seed, 41, 3 * N)).astype(np.float32)
    r, x = simulated(gpu, p)
    l1 = np.abs(x[:, 0::3]) + np.abs(x[:, 1::3]) + np.abs(x[:, 2::3])
    all_nan, some_nan = np.isnan(l1).all(axis=0), np.isnan(l1).any(axis=0)
    assert all_nan[N - 1] and np.any(some_nan & ~all_nan) and np.any(~some_nan)
    types, numbers = ground_truth(g, seed)
    got = r.calculate_final(types, numbers)
    want = oracle.calculate_final_f32(x, types, numbers)
    r.close()
    assert_bits(got, want)
    assert len(set(got["dice_score_over_threshold"].tolist())) > 10  # the sweep is not flat


def test_final_metrics_on_the_120k_voxel_shell(gpu):
    note("case: 64^3 ellipsoidal shell, 500 steps: device vs a vectorised count")
    p = synth.config1(num_steps=500).truth()
    types, numbers = ground_truth(p.geometry, 7)
    r, x = simulated(gpu, p)
    r.calculate_final(types, numbers)  # warm-up
    t0 = time.perf_counter()
    got = r.calculate_final(types, numbers)
    ms = 1e3 * (time.perf_counter() - t0)
    r.close()
    note(f"  N {p.num_voxels}, {types.size} cells: {ms:.2f} ms per call (with the uploads)")
    l1 = np.abs(x[:, 0::3]) + np.abs(x[:, 1::3]) + np.abs(x[:, 2::3])
    vmax = np.fmax.reduce(l1, axis=0)
    cell_max = np.full(types.size, np.nan, np.float32)
    cell_max[numbers >= 0] = vmax[numbers[numbers >= 0] // 3]
    truth = types == PATHOLOGICAL
    for i in range(101):
        with np.errstate(invalid="ignore"):
            pred = cell_max <= np.float32(i) / np.float32(100)
        tp, pp, gp = np.sum(pred & truth), np.sum(pred), np.sum(truth)
        fp, fn = pp - tp, gp - tp
        f = np.float32
        assert got["dice_score_over_threshold"][i] == (f(2 * tp) / f(2 * tp + fp + fn)
                                                      if 2 * tp + fp + fn else 0.0)
        assert got["iou_over_threshold"][i] == (f(tp) / f(tp + fp + fn) if tp + fp + fn else 0.0)
        assert got["precision_over_threshold"][i] == (f(tp) / f(pp) if pp else 0.0)
        assert got["recall_over_threshold"][i] == (f(tp) / f(gp) if gp else 1.0)


def test_final_metrics_edge_cases_and_errors(gpu):
    from cardiotrust_amd.gpu import CTError
    note("case: empty grid, a grid without voxels, bad voxel numbers")
    p = synth.make_problem(synth.cube(3), num_steps=20, max_delay=4, seed=8).truth()
    r, x = simulated(gpu, p)
    got = r.calculate_final(np.zeros(0, np.uint32), np.zeros(0, np.int32))
    assert_bits(got, oracle.calculate_final_f32(x, np.zeros(0, np.uint32), np.zeros(0, np.int32)))
    types = np.full(5, PATHOLOGICAL, np.uint32)
    none = np.full(5, -1, np.int32)
    assert_bits(r.calculate_final(types, none), oracle.calculate_final_f32(x, types, none))
    for bad in (4, 3 * p.num_voxels, -2):
        with pytest.raises(CTError):
            r.calculate_final(np.zeros(2, np.uint32), np.array([0, bad], np.int32))
    with pytest.raises(CTError):
        r.calculate_final(np.zeros(3, np.uint32), np.zeros(2, np.int32))
    r.close()
    # a scenario without voxels (N = 0): no maxima, nothing predicted Pathological
    g0 = synth.grid_geometry(np.zeros((2, 2, 2), dtype=bool), 2.5)
    p0 = synth.make_problem(g0, num_steps=10, max_delay=3, seed=9).truth()
    r0, x0 = simulated(gpu, p0)
    types = np.array([PATHOLOGICAL, VENTRICLE, 0, PATHOLOGICAL], np.uint32)
    none = np.full(4, -1, np.int32)
    got = r0.calculate_final(types, none)
    assert_bits(got, oracle.calculate_final_f32(x0, types, none))
    assert np.all(got["recall_over_threshold"] == 0.0)
    with pytest.raises(CTError):
        r0.calculate_final(types[:1], np.zeros(1, np.int32))  # no voxel 0 to point at
    r0.close()
