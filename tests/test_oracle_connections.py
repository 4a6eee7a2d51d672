"""The oracle's model connections (allpass.rs:212-289, allpass/delay.rs:14-111)
against the reference's own known answers and what the geometry fixes."""
import numpy as np
import pytest

from cardiotrust_amd import synth
from oracle import oracle


def grid_inputs(g, velocity=0.8):
    numbers = np.where(g.numbers >= 0, 3 * g.numbers, -1).astype(np.int32)
    idx = np.indices(g.shape).transpose(1, 2, 3, 0).astype(np.float64)
    pos = ((idx + 0.5) * g.voxel_size_mm).astype(np.float32)
    vel = np.full(g.shape, velocity, dtype=np.float32)
    return numbers, pos, vel


def test_delay_s_known_answers():
    """delay.rs:129-157: 1 m apart at 2 m/s is 0.5 s; (3, 4, 0) m is 2.5 s."""
    assert oracle.delay_s_f32([1000.0, 0, 0], [2000.0, 0, 0], 2.0) == pytest.approx(0.5, rel=1.2e-7)
    assert oracle.delay_s_f32([1000.0, 0, 0], [4000.0, 4000.0, 0], 2.0) == pytest.approx(2.5, rel=1.2e-7)


@pytest.mark.parametrize("which", ["blob", "shell"])
def test_output_state_indices_match_the_neighbour_form(which):
    g = synth.blob(4, 700) if which == "blob" else synth.ellipsoid_shell((20, 18, 16), (9, 8, 7), (4, 4, 3))
    numbers, pos, vel = grid_inputs(g)
    osi, dly = oracle.model_connections_f32(numbers, pos, vel, 2000.0, g.num_voxels)
    assert np.array_equal(osi, synth.output_state_indices(g.neighbours))
    assert np.all((dly > 0) == (g.neighbours >= 0))


def test_delays_face_edge_corner_and_symmetry():
    """A uniform grid: a connection with |offset|_1 = k has delay
    size/1000/v * rate * sqrt(k) (the reference's max check, delay.rs:160-180,
    is the k = 2 case of its default model); the reverse connection has the
    same delay bit for bit."""
    g = synth.cube(5)
    numbers, pos, vel = grid_inputs(g, velocity=0.5)
    osi, dly = oracle.model_connections_f32(numbers, pos, vel, 2000.0, g.num_voxels)
    base = (2.5 / 1000.0) / 0.5 * 2000.0
    for d in range(26):
        k = int(np.abs(synth.OFFSETS[d]).sum())
        have = g.neighbours[:, d] >= 0
        assert np.allclose(dly[have, d], base * np.sqrt(k), rtol=3e-7)
        w = g.neighbours[have, d]
        assert np.array_equal(dly[w, 25 - d].view(np.uint32), dly[have, d].view(np.uint32))
    assert dly.max() == pytest.approx(base * np.sqrt(3.0), rel=3e-7)


def test_sub_sample_delay_is_an_error():
    g = synth.cube(3, voxel_size_mm=0.1)
    numbers, pos, vel = grid_inputs(g, velocity=1.0)
    with pytest.raises(ValueError):
        oracle.model_connections_f32(numbers, pos, vel, 2000.0, g.num_voxels)
