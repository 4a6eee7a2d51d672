"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle_ct.so (built by oracle/Makefile) and runs the
reference-faithful CPU epoch (`algorithm.rs:124-281`) on numpy arrays in the
reference's layouts, in float32 (operation-order faithful) or float64 (the
error yardstick). Only tests, the smoke check and the benchmark's CPU-baseline
leg import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.environ.get("CT_ORACLE_LIB") or os.path.join(_HERE, "liboracle_ct.so")
_lib = None


class _Problem(ctypes.Structure):
    _fields_ = [
        ("num_voxels", ctypes.c_int32),
        ("num_sensors", ctypes.c_int32),
        ("num_steps", ctypes.c_int32),
        ("num_beats", ctypes.c_int32),
    ] + [(name, ctypes.c_void_p) for name in (
        "gains", "output_state_indices", "coefs", "delays", "initial_delays",
        "control_matrix", "control_function", "measurement_matrix", "actual_measurements",
        "ap_outputs_now", "ap_outputs_last", "system_states", "measurements", "residuals",
        "average_delays", "average_delays_valid",
        "d_gains", "d_coefs", "d_coefs_iir", "d_coefs_fir", "mapped_residuals",
        "maximum_regularization", "maximum_regularization_sum", "d_step",
        "gains_first_moment", "gains_second_moment", "coefs_first_moment", "coefs_second_moment",
        "loss", "loss_mse", "loss_maximum_regularization",
        "loss_batch", "loss_mse_batch", "loss_maximum_regularization_batch")] + [
        ("num_batch_entries", ctypes.c_int32)]


class _Algorithm(ctypes.Structure):
    _fields_ = [
        ("learning_rate", ctypes.c_float),
        ("batch_size", ctypes.c_int32),
        ("optimizer", ctypes.c_int32),
        ("mse_strength", ctypes.c_float),
        ("maximum_regularization_strength", ctypes.c_float),
        ("maximum_regularization_threshold", ctypes.c_float),
        ("difference_regularization_strength", ctypes.c_float),
        ("smoothness_regularization_strength", ctypes.c_float),
        ("freeze_gains", ctypes.c_int32),
        ("freeze_delays", ctypes.c_int32),
        ("ap_derivative", ctypes.c_int32),
    ]


def build() -> str:
    """Compile liboracle_ct.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        for name in ("ct_oracle_run_epoch_f32", "ct_oracle_run_epoch_f64"):
            f = getattr(L, name)
            f.argtypes = [P(_Problem), P(_Algorithm), ctypes.c_void_p, P(ctypes.c_int32)]
            f.restype = ctypes.c_int
        for name in ("ct_oracle_simulate_f32", "ct_oracle_simulate_f64"):
            f = getattr(L, name)
            f.argtypes = [P(_Problem)]
            f.restype = ctypes.c_int
        for name in ("ct_oracle_adjoint_f64", "ct_oracle_adjoint_f32"):
            f = getattr(L, name)
            f.argtypes = [P(_Problem), P(_Algorithm), ctypes.c_int32, P(ctypes.c_double),
                          ctypes.c_void_p, ctypes.c_void_p]
            f.restype = ctypes.c_int
        L.ct_oracle_offset_to_gain_index.argtypes = [ctypes.c_int32] * 4
        L.ct_oracle_offset_to_gain_index.restype = ctypes.c_int32
        L.ct_oracle_gain_index_to_offset.argtypes = [ctypes.c_int32, P(ctypes.c_int32)]
        L.ct_oracle_gain_index_to_offset.restype = ctypes.c_int32
        L.ct_oracle_offset_to_delay_index.argtypes = [ctypes.c_int32] * 3
        L.ct_oracle_offset_to_delay_index.restype = ctypes.c_int32
        L.ct_oracle_delay_index_to_offset.argtypes = [ctypes.c_int32, P(ctypes.c_int32)]
        L.ct_oracle_delay_index_to_offset.restype = ctypes.c_int32
        L.ct_oracle_from_samples_to_coef.argtypes = [ctypes.c_float]
        L.ct_oracle_from_samples_to_coef.restype = ctypes.c_float
        L.ct_oracle_from_samples_to_usize.argtypes = [ctypes.c_float]
        L.ct_oracle_from_samples_to_usize.restype = ctypes.c_int64
        L.ct_oracle_from_coef_to_samples.argtypes = [ctypes.c_float]
        L.ct_oracle_from_coef_to_samples.restype = ctypes.c_float
        L.ct_oracle_average_delays_f32.argtypes = [ctypes.c_int32] + [ctypes.c_void_p] * 5
        L.ct_oracle_average_delays_f32.restype = ctypes.c_int
        L.ct_oracle_update_sgd_f32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                               ctypes.c_float, ctypes.c_int64]
        L.ct_oracle_update_sgd_f32.restype = None
        L.ct_oracle_roll_delays_f32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        L.ct_oracle_roll_delays_f32.restype = None
        L.ct_oracle_mapped_residuals_f32.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_int32, ctypes.c_int32,
                                                     ctypes.c_void_p]
        L.ct_oracle_mapped_residuals_f32.restype = None
        L.ct_oracle_unrolled_dot_f32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        L.ct_oracle_unrolled_dot_f32.restype = ctypes.c_float
        L.ct_oracle_unrolled_sum_f32.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.ct_oracle_unrolled_sum_f32.restype = ctypes.c_float
        L.ct_oracle_measurement_matrix_f32.argtypes = [
            ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_int32, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p]
        L.ct_oracle_measurement_matrix_f32.restype = None
        L.ct_oracle_model_connections_f32.argtypes = [
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_float, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
        L.ct_oracle_model_connections_f32.restype = ctypes.c_int
        L.ct_oracle_spherical_f32.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_float] + [ctypes.c_void_p] * 7
        L.ct_oracle_spherical_f32.restype = ctypes.c_int
        L.ct_oracle_calculate_final_f32.argtypes = [
            ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_int64] + [ctypes.c_void_p] * 4
        L.ct_oracle_calculate_final_f32.restype = None
        L.ct_oracle_delay_s_f32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float]
        L.ct_oracle_delay_s_f32.restype = ctypes.c_float
        _lib = L
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _algorithm(config) -> _Algorithm:
    return _Algorithm(float(config.learning_rate), int(config.batch_size), int(config.optimizer),
                      float(config.mse_strength), float(config.maximum_regularization_strength),
                      float(config.maximum_regularization_threshold),
                      float(config.difference_regularization_strength),
                      float(config.smoothness_regularization_strength),
                      int(bool(config.freeze_gains)), int(bool(config.freeze_delays)),
                      int(config.ap_derivative))


class OracleResults:
    """Host state of one scenario (model, estimations, derivatives, metrics)
    mirroring the reference `Results` (`scenario/results.rs:28-34`)."""

    def __init__(self, problem, dtype=np.float32, optimizer: int = 0,
                 num_batch_entries: int = 1):
        self.dtype = np.dtype(dtype)
        R = self.dtype
        p = problem
        N, S, T, B = p.num_voxels, p.num_sensors, p.num_steps, p.num_beats
        n = 3 * N
        self.N, self.S, self.T, self.B = N, S, T, B
        c = lambda a, dt=R: np.ascontiguousarray(np.asarray(a).astype(dt, copy=True))
        self.gains = c(p.gains)
        self.output_state_indices = c(p.output_state_indices, np.int32)
        self.coefs = c(p.coefs)
        self.delays = c(p.delays, np.int32)
        self.initial_delays = c(p.initial_delays)
        self.control_matrix = c(p.control_matrix)
        self.control_function = c(p.control_function)
        self.measurement_matrix = c(p.measurement_matrix)
        meas = p.measurements if p.measurements is not None else np.zeros((B, T, S))
        self.actual_measurements = c(meas)
        z = lambda *shape: np.zeros(shape, dtype=R)
        self.ap_outputs_now = z(n, 78)
        self.ap_outputs_last = z(n, 78)
        self.system_states = z(T, n)
        self.measurements = z(B, T, S)
        self.residuals = z(S)
        self.average_delays = z(N)
        self.average_delays_valid = np.zeros(N, dtype=np.uint8)
        self.d_gains = z(n, 78)
        self.d_coefs = z(N, 26)
        self.d_coefs_iir = z(n, 78)
        self.d_coefs_fir = z(n, 78)
        self.mapped_residuals = z(n)
        self.maximum_regularization = z(n)
        self.maximum_regularization_sum = z(1)
        self.d_step = np.ones(1, dtype=np.int64)
        adam = optimizer == 1
        self.gains_first_moment = z(n, 78) if adam else None
        self.gains_second_moment = z(n, 78) if adam else None
        self.coefs_first_moment = z(N, 26) if adam else None
        self.coefs_second_moment = z(N, 26) if adam else None
        self.loss = z(T)
        self.loss_mse = z(T)
        self.loss_maximum_regularization = z(T)
        self.loss_batch = z(num_batch_entries)
        self.loss_mse_batch = z(num_batch_entries)
        self.loss_maximum_regularization_batch = z(num_batch_entries)
        self.batch_index = ctypes.c_int32(0)

    def _struct(self) -> _Problem:
        s = _Problem()
        s.num_voxels, s.num_sensors, s.num_steps, s.num_beats = self.N, self.S, self.T, self.B
        for name, _ in _Problem._fields_[4:-1]:
            setattr(s, name, _ptr(getattr(self, name)))
        s.num_batch_entries = self.loss_batch.shape[0]
        return s

    def run_epoch(self, config, beat_order=None) -> None:
        """One epoch (`algorithm.rs:124-281`); `beat_order` replaces the
        reference's unseeded shuffle."""
        a = _algorithm(config)
        order = None
        if beat_order is not None:
            order = np.ascontiguousarray(beat_order, dtype=np.int32)
        s = self._struct()
        f = lib().ct_oracle_run_epoch_f64 if self.dtype == np.float64 else lib().ct_oracle_run_epoch_f32
        rc = f(ctypes.byref(s), ctypes.byref(a), _ptr(order), ctypes.byref(self.batch_index))
        if rc != 0:
            raise RuntimeError(f"oracle run_epoch failed ({rc})")

    def adjoint(self, config, beat: int = 0):
        """Objective J and its exact gradient (dJ/dgains (3N,78), dJ/dcoefs
        (N,26)) for one beat, by reverse mode (see oracle.h)."""
        a = _algorithm(config)
        J = ctypes.c_double()
        dG = np.zeros((3 * self.N, 78), dtype=self.dtype)
        dA = np.zeros((self.N, 26), dtype=self.dtype)
        s = self._struct()
        f = lib().ct_oracle_adjoint_f64 if self.dtype == np.float64 else lib().ct_oracle_adjoint_f32
        rc = f(ctypes.byref(s), ctypes.byref(a), int(beat), ctypes.byref(J), _ptr(dG), _ptr(dA))
        if rc != 0:
            raise RuntimeError(f"oracle adjoint failed ({rc})")
        return J.value, dG, dA

    def simulate(self) -> np.ndarray:
        """Forward-only pass over every beat; returns measurements (B,T,S)."""
        s = self._struct()
        f = lib().ct_oracle_simulate_f64 if self.dtype == np.float64 else lib().ct_oracle_simulate_f32
        rc = f(ctypes.byref(s))
        if rc != 0:
            raise RuntimeError(f"oracle simulate failed ({rc})")
        return self.measurements


def simulate_f32(problem) -> np.ndarray:
    """Noise-free measurements of `problem`'s parameters (float32 oracle)."""
    return OracleResults(problem, np.float32).simulate().copy()


def measurement_matrix_f32(voxel_positions_mm, sensor_positions_mm, sensor_orientations,
                           sensor_offsets_mm, voxel_size_mm):
    """(B, S, 3N) Biot-Savart matrix in the reference's f32 operation order
    (measurement.rs:66-123); inputs are cast to float32 as the reference holds
    them."""
    vp = np.ascontiguousarray(voxel_positions_mm, dtype=np.float32).reshape(-1, 3)
    sp = np.ascontiguousarray(sensor_positions_mm, dtype=np.float32).reshape(-1, 3)
    so = np.ascontiguousarray(sensor_orientations, dtype=np.float32).reshape(-1, 3)
    off = np.ascontiguousarray(sensor_offsets_mm, dtype=np.float32).reshape(-1, 3)
    N, S, B = vp.shape[0], sp.shape[0], off.shape[0]
    out = np.zeros((B, S, 3 * N), dtype=np.float32)
    lib().ct_oracle_measurement_matrix_f32(N, vp.ctypes.data, S, sp.ctypes.data, so.ctypes.data,
                                           B, off.ctypes.data, float(voxel_size_mm),
                                           out.ctypes.data)
    return out


def delay_s_f32(input_position_mm, output_position_mm, velocity_m_per_s):
    a = np.ascontiguousarray(input_position_mm, dtype=np.float32)
    b = np.ascontiguousarray(output_position_mm, dtype=np.float32)
    return float(lib().ct_oracle_delay_s_f32(a.ctypes.data, b.ctypes.data, float(velocity_m_per_s)))


def model_connections_f32(numbers, positions_mm, velocities_m_per_s, sample_rate_hz, num_voxels):
    """(output_state_indices (3N, 78) int32, delay_samples (N, 26) float32) of a
    grid (nx, ny, nz): numbers = first state index per cell or -1; raises
    ValueError where the reference errors (a delay below one sample)."""
    num = np.ascontiguousarray(numbers, dtype=np.int32)
    nx, ny, nz = num.shape
    pos = np.ascontiguousarray(positions_mm, dtype=np.float32).reshape(nx, ny, nz, 3)
    vel = np.ascontiguousarray(velocities_m_per_s, dtype=np.float32).reshape(nx, ny, nz)
    osi = np.empty((3 * num_voxels, 78), dtype=np.int32)
    dly = np.empty((num_voxels, 26), dtype=np.float32)
    rc = lib().ct_oracle_model_connections_f32(nx, ny, nz, num.ctypes.data, pos.ctypes.data,
                                               vel.ctypes.data, float(sample_rate_hz), num_voxels,
                                               osi.ctypes.data, dly.ctypes.data)
    if rc != 0:
        raise ValueError("a delay is below one sample")
    return osi, dly


FINAL_METRICS = ("dice_score_over_threshold", "iou_over_threshold",
                 "precision_over_threshold", "recall_over_threshold")


def calculate_final_f32(x, voxel_types, voxel_numbers):
    """metrics::calculate_final over states x (T, n): dict of the four
    (101,) arrays named as the reference's Metrics fields. voxel_types:
    VoxelType ordinals per grid cell (Pathological = 6); voxel_numbers: each
    cell's first state index, -1 for None (the reference's .npy forms)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    types = np.ascontiguousarray(voxel_types, dtype=np.uint32).ravel()
    numbers = np.ascontiguousarray(voxel_numbers, dtype=np.int32).ravel()
    if types.shape != numbers.shape:
        raise ValueError("voxel_types and voxel_numbers differ in shape")
    T, n = x.shape
    real = numbers[numbers >= 0]
    if (numbers.size and numbers.min() < -1) or (real.size and (real.max() + 2 >= n
                                                                or np.any(real % 3))):
        raise ValueError("voxel numbers must be -1 or multiples of 3 below n")
    out = {k: np.zeros(101, np.float32) for k in FINAL_METRICS}
    lib().ct_oracle_calculate_final_f32(x.ctypes.data, T, n, types.ctypes.data,
                                        numbers.ctypes.data, types.size,
                                        *(out[k].ctypes.data for k in FINAL_METRICS))
    return out


def spherical_f32(x, sample_rate_hz):
    """dict of magnitude/theta/phi (T, N), their values at each voxel's
    maximum magnitude and activation_time_ms (N) for states x (T, 3N);
    ValueError where the reference errors."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    T, n = x.shape
    N = n // 3
    out = {k: np.zeros((T, N), np.float32) for k in ("magnitude", "theta", "phi")}
    out.update({k: np.zeros(N, np.float32) for k in ("magnitude_max", "theta_max", "phi_max",
                                                     "activation_time_ms")})
    rc = lib().ct_oracle_spherical_f32(
        x.ctypes.data, T, N, float(sample_rate_hz), out["magnitude"].ctypes.data,
        out["theta"].ctypes.data, out["phi"].ctypes.data, out["magnitude_max"].ctypes.data,
        out["theta_max"].ctypes.data, out["phi_max"].ctypes.data,
        out["activation_time_ms"].ctypes.data)
    if rc != 0:
        raise ValueError("a voxel has no non-NaN magnitude")
    return out
