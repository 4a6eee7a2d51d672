"""The oracle's metrics::calculate_final (algorithm/metrics.rs:225-441).

Pinned three ways: a hand-computed case (maxima that sit exactly on a
threshold, a cell without a voxel, a ground-truth Pathological cell that
holds no voxel); the reference's values for empty denominators; and, on
random grids with NaN and infinite states, scikit-learn's f1 / jaccard /
precision / recall scores of the same predictions (the Dice score is the
binary F1 score, the IoU the Jaccard index), with the reference's
zero-division values.
"""
import numpy as np
import pytest

from oracle import oracle

PATHOLOGICAL, VENTRICLE, NONE = 6, 5, 0


def grid(mask):
    """voxel_numbers in the reference's form: 3 per voxel in C order, -1 for None
    (VoxelNumbers::from_voxel_types, model/spatial/voxels.rs:407-423)."""
    mask = np.asarray(mask, bool).ravel()
    numbers = np.full(mask.size, -1, np.int32)
    numbers[mask] = 3 * np.arange(mask.sum(), dtype=np.int32)
    return numbers


def states_with_maxima(maxima, T=5):
    """x (T, 3N) whose L1 maxima over the steps are `maxima` (first component)."""
    N = len(maxima)
    x = np.zeros((T, 3 * N), np.float32)
    for v, m in enumerate(maxima):
        x[T // 2, 3 * v] = np.float32(m)
        x[0, 3 * v + 1] = np.float32(m) * np.float32(0.25)
    return x


def test_known_answer():
    # cells: 4 voxels, then a cell without a voxel that is Pathological in the
    # ground truth (always a false negative)
    numbers = np.array([0, 3, 6, 9, -1], np.int32)
    types = np.array([PATHOLOGICAL, PATHOLOGICAL, VENTRICLE, VENTRICLE, PATHOLOGICAL], np.uint32)
    x = states_with_maxima([0.1, 0.5, 0.5, 2.0])
    out = oracle.calculate_final_f32(x, types, numbers)
    d, j, p, r = (out[k] for k in oracle.FINAL_METRICS)
    # threshold 0.05: nothing predicted; tp 0, gt positives 3
    assert (d[5], j[5], p[5], r[5]) == (0.0, 0.0, 0.0, 0.0)
    # 0.1f equals 10 / 100 in f32, so voxel 0 is predicted from i = 10 on
    assert r[9] == 0.0 and r[10] == np.float32(1) / np.float32(3)
    # i = 30: predicted {0}: tp 1, fp 0, fn 2
    assert d[30] == np.float32(2) / np.float32(4)
    assert j[30] == np.float32(1) / np.float32(3)
    assert p[30] == 1.0 and r[30] == np.float32(1) / np.float32(3)
    # i = 50: predicted {0, 1, 2}: tp 2, fp 1, fn 1
    assert d[50] == np.float32(4) / np.float32(6)
    assert j[50] == np.float32(2) / np.float32(4)
    assert p[50] == np.float32(2) / np.float32(3)
    assert r[50] == np.float32(2) / np.float32(3)
    # i = 100: voxel 3 (max 2) still not predicted
    assert p[100] == np.float32(2) / np.float32(3)


def test_empty_denominators():
    x = states_with_maxima([0.3, 0.7])
    # no ground-truth positives: recall 1; nothing predicted at i = 0: precision 0;
    # no positives at all at i = 0: Dice and IoU 0
    out = oracle.calculate_final_f32(x, np.array([VENTRICLE, VENTRICLE], np.uint32),
                                     np.array([0, 3], np.int32))
    assert np.all(out["recall_over_threshold"] == 1.0)
    assert out["precision_over_threshold"][0] == 0.0
    assert out["dice_score_over_threshold"][0] == 0.0 and out["iou_over_threshold"][0] == 0.0
    # from i = 30 on voxel 0 is a false positive: Dice, IoU and precision 0
    assert out["dice_score_over_threshold"][30] == 0.0
    assert out["precision_over_threshold"][30] == 0.0
    # no cells at all
    out = oracle.calculate_final_f32(x, np.zeros(0, np.uint32), np.zeros(0, np.int32))
    assert np.all(out["recall_over_threshold"] == 1.0)
    for k in oracle.FINAL_METRICS[:3]:
        assert np.all(out[k] == 0.0)


def test_all_nan_voxel_is_never_pathological_and_nan_steps_are_skipped():
    x = states_with_maxima([0.2, 0.2, 0.2], T=6)
    x[:, 0:3] = np.nan  # voxel 0: every step NaN
    x[1, 4] = np.nan    # voxel 1: one NaN step, maximum 0.2 from the others
    out = oracle.calculate_final_f32(x, np.full(3, PATHOLOGICAL, np.uint32),
                                     np.array([0, 3, 6], np.int32))
    # from i = 20 on voxels 1 and 2 are predicted, voxel 0 never
    assert out["recall_over_threshold"][19] == 0.0
    assert out["recall_over_threshold"][20] == np.float32(2) / np.float32(3)
    assert out["recall_over_threshold"][100] == np.float32(2) / np.float32(3)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_against_scikit_learn(seed):
    skm = pytest.importorskip("sklearn.metrics")
    rng = np.random.default_rng(seed)
    shape = (6, 5, 4)
    mask = rng.random(shape) < 0.7
    numbers = grid(mask)
    N = int(mask.sum())
    T = 9
    x = (rng.standard_normal((T, 3 * N)) * rng.uniform(0.05, 0.5, 3 * N)).astype(np.float32)
    nan_cells = rng.random((T, 3 * N)) < 0.05
    x[nan_cells] = np.nan
    x[:, 0:3] = np.nan                # an all-NaN voxel
    x[2, 3] = np.inf                  # a voxel with an infinite state
    types = np.where(rng.random(mask.size) < 0.4, PATHOLOGICAL,
                     np.where(mask.ravel(), VENTRICLE, NONE)).astype(np.uint32)
    out = oracle.calculate_final_f32(x, types, numbers)

    l1 = np.abs(x[:, 0::3]) + np.abs(x[:, 1::3]) + np.abs(x[:, 2::3])
    with np.errstate(invalid="ignore"):
        vmax = np.fmax.reduce(l1, axis=0)  # NaN only where every step is NaN
    cell_max = np.full(mask.size, np.nan, np.float32)
    cell_max[numbers >= 0] = vmax[numbers[numbers >= 0] // 3]
    truth = types == PATHOLOGICAL
    for i in range(101):
        thr = np.float32(i) / np.float32(100)
        with np.errstate(invalid="ignore"):
            pred = cell_max <= thr
        want = (skm.f1_score(truth, pred, zero_division=0.0),
                skm.jaccard_score(truth, pred, zero_division=0.0),
                skm.precision_score(truth, pred, zero_division=0.0),
                skm.recall_score(truth, pred, zero_division=1.0))
        for k, w in zip(oracle.FINAL_METRICS, want):
            assert out[k][i] == np.float32(w), (i, k, out[k][i], w)
