"""The oracle against the reference's own known-answer tests.

Each test restates the inputs and expected values of a reference unit test
(cited file:line, relative to the reference's src/core/).
"""
import ctypes

import numpy as np
import pytest

from oracle import oracle


def test_offset_to_gain_index_golden(oracle_lib):
    # model/functional/allpass.rs:697-730
    cases = [((-1, -1, -1, 2), 2), ((-1, -1, 0, 2), 5), ((-1, -1, 1, 2), 8),
             ((1, 1, 1, 2), 77), ((0, 1, -1, 0), 42), ((0, 1, 0, 0), 45),
             ((1, 0, -1, 0), 60), ((1, 0, 0, 0), 63)]
    for args, want in cases:
        assert oracle_lib.ct_oracle_offset_to_gain_index(*args) == want
    assert oracle_lib.ct_oracle_offset_to_gain_index(0, 0, 0, 1) == -1


def test_gain_and_delay_index_round_trip(oracle_lib):
    # allpass.rs:532-618: the four helpers are mutual inverses on their domains
    out4 = (ctypes.c_int32 * 4)()
    out3 = (ctypes.c_int32 * 3)()
    seen_g, seen_d = set(), set()
    for x in (-1, 0, 1):
        for y in (-1, 0, 1):
            for z in (-1, 0, 1):
                if (x, y, z) == (0, 0, 0):
                    continue
                d = oracle_lib.ct_oracle_offset_to_delay_index(x, y, z)
                assert 0 <= d < 26
                seen_d.add(d)
                assert oracle_lib.ct_oracle_delay_index_to_offset(d, out3) == 0
                assert tuple(out3) == (x, y, z)
                for o in range(3):
                    g = oracle_lib.ct_oracle_offset_to_gain_index(x, y, z, o)
                    assert g == 3 * d + o  # gain column = 3 * connection + out
                    seen_g.add(g)
                    assert oracle_lib.ct_oracle_gain_index_to_offset(g, out4) == 0
                    assert tuple(out4) == (x, y, z, o)
    assert seen_d == set(range(26)) and seen_g == set(range(78))
    assert oracle_lib.ct_oracle_gain_index_to_offset(78, out4) == -1


def test_from_samples_golden(oracle_lib):
    # allpass.rs:678-695
    for s, want in [(1.0, 1), (1.2, 1), (10.9, 10), (10.0, 10)]:
        assert oracle_lib.ct_oracle_from_samples_to_usize(s) == want
    for s in (0.5, 1.5, 99999.5):
        assert oracle_lib.ct_oracle_from_samples_to_coef(s) == pytest.approx(1.0 / 3.0, rel=1e-6)
    for s in (0.0, 1.0, 99999.0):
        assert oracle_lib.ct_oracle_from_samples_to_coef(s) == pytest.approx(0.9999, rel=1e-6)
    # from_coef_to_samples inverts from_samples_to_coef on the fractional part
    for frac in (0.1, 0.25, 0.5, 0.75):
        a = oracle_lib.ct_oracle_from_samples_to_coef(3.0 + frac)
        assert oracle_lib.ct_oracle_from_coef_to_samples(a) == pytest.approx(frac, abs=1e-6)


def _average_delays(gains, coefs, delays):
    N = coefs.shape[0]
    avg = np.zeros(N, dtype=np.float32)
    valid = np.zeros(N, dtype=np.uint8)
    g = np.ascontiguousarray(gains, dtype=np.float32)
    c = np.ascontiguousarray(coefs, dtype=np.float32)
    d = np.ascontiguousarray(delays, dtype=np.int32)
    oracle.lib().ct_oracle_average_delays_f32(N, g.ctypes.data, c.ctypes.data, d.ctypes.data,
                                              avg.ctypes.data, valid.ctypes.data)
    return avg, valid


def test_average_delays_golden(oracle_lib):
    L = oracle_lib
    # refinement/derivation.rs:868-888
    avg, valid = _average_delays(np.ones((3, 78)), np.full((1, 26), L.ct_oracle_from_samples_to_coef(0.5)),
                                 np.full((1, 26), 2))
    assert valid[0] and abs(avg[0] - 1.836_931_7) < 1e-6
    # derivation.rs:890-915
    avg, valid = _average_delays(np.ones((6, 78)), np.full((2, 26), L.ct_oracle_from_samples_to_coef(0.4)),
                                 np.full((2, 26), 2))
    assert abs(avg[0] - 1.763_453_2) < 1e-4 and abs(avg[1] - 1.763_453) < 1e-4
    # derivation.rs:917-933
    avg, valid = _average_delays(np.zeros((3, 78)), np.full((1, 26), L.ct_oracle_from_samples_to_coef(0.5)),
                                 np.full((1, 26), 2))
    assert not valid[0]
    # derivation.rs:935-958
    g = np.zeros((3, 78))
    g[0, 10], g[1, 20], g[2, 30] = 1.0, 4.0, 2.0
    avg, valid = _average_delays(g, np.full((1, 26), L.ct_oracle_from_samples_to_coef(0.1)),
                                 np.full((1, 26), 2))
    assert valid[0] and abs(avg[0] - 1.504_952_5) < 1e-6


def test_update_golden(oracle_lib):
    # refinement/update.rs:235-261: derivatives -0.5, lr 1, batch 1 -> values +0.5
    g = np.zeros(10 * 78, dtype=np.float32)
    d = np.full_like(g, -0.5)
    oracle_lib.ct_oracle_update_sgd_f32(g.ctypes.data, d.ctypes.data, g.size, 1.0, 1)
    assert np.array_equal(g, -d)
    a = np.zeros(4 * 26, dtype=np.float32)
    D = np.zeros(4 * 26, dtype=np.int32)
    da = np.full_like(a, -0.5)
    oracle_lib.ct_oracle_update_sgd_f32(a.ctypes.data, da.ctypes.data, a.size, 1.0, 1)
    oracle_lib.ct_oracle_roll_delays_f32(a.ctypes.data, D.ctypes.data, a.size)
    assert np.array_equal(a, -da) and np.all(D == 0)


def test_roll_delays_all_branches(oracle_lib):
    # update.rs:206-228, every branch by brute force
    a = np.array([0.99995, 0.99995, 0.00005, 0.00005, 0.5, 0.9999, 0.0001], dtype=np.float32)
    D = np.array([5, 1, 5, 1000, 7, 3, 3], dtype=np.int32)
    oracle_lib.ct_oracle_roll_delays_f32(a.ctypes.data, D.ctypes.data, a.size)
    want_a = np.array([2e-4, 1 - 1e-4, np.float32(1) - np.float32(2e-4), 1e-4, 0.5, 0.9999, 0.0001],
                      dtype=np.float32)
    assert np.array_equal(D, [4, 1, 6, 1000, 7, 3, 3])
    assert np.allclose(a, want_a, rtol=0, atol=1e-7)


def test_mapped_residuals_closed_form(oracle_lib):
    # algorithm/gpu/derivation.rs:552-660: 4 sensors x 2 states
    H = np.array([[1.0, 0.0], [0.0, 1.0], [1.0, 1.0], [2.0, -1.0]], dtype=np.float32)
    r = np.array([0.5, -0.3, 0.8, 0.2], dtype=np.float32)
    e = np.zeros(2, dtype=np.float32)
    oracle_lib.ct_oracle_mapped_residuals_f32(H.ctypes.data, r.ctypes.data, 4, 2, e.ctypes.data)
    assert np.allclose(e, [1.0 * 0.5 + 0.8 + 2.0 * 0.2, -0.3 + 0.8 - 0.2], atol=1e-6)


def test_unrolled_reductions_structure(oracle_lib):
    # ndarray unrolled_dot / unrolled_fold: exact on integers, and the 8-way
    # interleaving is visible on a sequence that cancels catastrophically.
    for n in (0, 1, 7, 8, 9, 15, 16, 17, 100):
        a = np.arange(1, n + 1, dtype=np.float32)
        b = np.ones(n, dtype=np.float32)
        assert oracle_lib.ct_oracle_unrolled_dot_f32(a.ctypes.data, b.ctypes.data, n) == n * (n + 1) / 2
        assert oracle_lib.ct_oracle_unrolled_sum_f32(a.ctypes.data, n) == n * (n + 1) / 2
    # [1e8, 1, 0*6, -1e8, ...]: sequential left-to-right gives 0; the
    # interleaved sum keeps the 1 in its own partial (p1) and returns 1.
    x = np.zeros(16, dtype=np.float32)
    x[0], x[1], x[8] = 1e8, 1.0, -1e8
    assert oracle_lib.ct_oracle_unrolled_sum_f32(x.ctypes.data, 16) == 1.0
    assert np.float32(np.float32(np.float32(1e8) + 1) - 1e8) == 0.0
