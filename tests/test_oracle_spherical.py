"""The oracle's spherical post-processing (data/shapes.rs:150-330) against
closed forms and argmax_skipnan's rules (first maximum, NaN skipped, error
when nothing is left)."""
import numpy as np
import pytest

from oracle import oracle


def test_spherical_closed_forms():
    x = np.array([[0.0, 0.0, 2.0, 3.0, 0.0, 0.0, 0.0, -1.0, 0.0, 1.0, 1.0, 0.0]], np.float32)
    o = oracle.spherical_f32(x, 2000.0)
    assert np.array_equal(o["magnitude"][0], np.float32([2.0, 3.0, 1.0, 2.0]))
    assert np.allclose(o["theta"][0], [0.0, np.pi / 2, np.pi / 2, np.pi / 2], atol=1e-7)
    assert np.allclose(o["phi"][0], [0.0, 0.0, -np.pi / 2, np.pi / 4], atol=1e-7)


def test_first_maximum_nan_skipped_and_activation_time():
    mags = np.array([[1.0, np.nan, 0.0], [3.0, 1.0, 0.0], [3.0, 2.0, 0.0], [2.0, np.nan, 0.0]],
                    np.float32)
    x = np.zeros((4, 9), np.float32)
    x[:, 0::3] = mags  # all of the magnitude on the first component
    o = oracle.spherical_f32(x, 500.0)
    assert np.array_equal(o["activation_time_ms"], np.float32([2.0, 4.0, 0.0]))  # index / 500 * 1000
    assert np.array_equal(o["magnitude_max"], np.float32([3.0, 2.0, 0.0]))
    assert np.isnan(o["theta_max"][2])  # acos(0 / 0) at the all-zero voxel, as the reference


def test_all_nan_voxel_is_an_error():
    x = np.full((3, 3), np.nan, np.float32)
    with pytest.raises(ValueError):
        oracle.spherical_f32(x, 1000.0)
