"""The oracle's Biot-Savart matrix (measurement.rs:66-123 in the reference's
f32 operations) against what physics fixes and against the FP64 host builder.
The reference's own test (equality_sparse_full, measurement.rs:452-483) only
compares two sensor layouts with each other."""
import numpy as np

from cardiotrust_amd import synth
from oracle import oracle

MU0 = 1.25663706212e-6


def common(size_mm):
    return MU0 * (size_mm / 1000.0) ** 3 / (4.0 * np.pi) * 1e12


def test_current_element_right_hand_rule():
    """A current element along x at the origin and a sensor z mm above it:
    B = mu0/(4 pi) (J x d) / |d|^3 points along -y with |B| = c / z^2; along x
    and z the sensor reads 0. A current along y gives +x. Swapped cross-product
    indices, a sign error or a wrong power of |d| fail these values."""
    z = 80.0
    vox = np.zeros((1, 3))
    sens = np.array([[0.0, 0.0, z]] * 3)
    ori = np.eye(3)  # sensors measuring along x, y, z
    H = oracle.measurement_matrix_f32(vox, sens, ori, np.zeros((1, 3)), 2.5)[0]
    c = common(2.5) / (z / 1000.0) ** 2
    want = np.array([[0.0, c, 0.0],    # x-sensor: J_y gives B_x = +c
                     [-c, 0.0, 0.0],   # y-sensor: J_x gives B_y = -c
                     [0.0, 0.0, 0.0]])  # z-sensor: nothing along the line of sight
    assert np.allclose(H, want, rtol=1e-6, atol=1e-12 * c)


def test_inverse_square_and_beat_offsets():
    """|H| falls as 1/|d|^2 along the axis, and a beat offset equals moving the
    sensors by that offset."""
    vox = np.zeros((1, 3))
    ori = np.array([[0.0, 1.0, 0.0]])
    h = [oracle.measurement_matrix_f32(vox, np.array([[0.0, 0.0, z]]), ori, np.zeros((1, 3)),
                                       2.5)[0, 0, 0] for z in (50.0, 100.0, 200.0)]
    assert np.allclose([h[0] / h[1], h[1] / h[2]], [4.0, 4.0], rtol=1e-6)
    g = synth.blob(7, 200)
    s = synth.sensor_grid(g)
    off = np.array([[0.0, 0.0, 0.0], [7.5, -2.5, 5.0]], dtype=np.float32)
    H = oracle.measurement_matrix_f32(g.positions_mm, s.positions_mm, s.orientations, off, 2.5)
    moved = (s.positions_mm.astype(np.float32) + off[1]).astype(np.float32)
    H1 = oracle.measurement_matrix_f32(g.positions_mm, moved, s.orientations, off[:1], 2.5)
    assert np.array_equal(H[1], H1[0])


def test_oracle_matches_the_fp64_builder():
    g = synth.blob(3, 300)
    s = synth.sensor_grid(g, num_beats=2, beat_offsets_mm=np.array([[0, 0, 0], [10.0, -5, 3]]))
    H32 = oracle.measurement_matrix_f32(g.positions_mm, s.positions_mm, s.orientations,
                                        s.beat_offsets_mm, g.voxel_size_mm)
    H64 = synth.measurement_matrix(g, s)
    assert H32.shape == H64.shape == (2, 108, 900)
    assert np.max(np.abs(H32 - H64)) <= 1e-6 * np.max(np.abs(H64))
