"""Python mirror of the reference's GPU epoch API over the C ABI.

Names follow `src/core/algorithm/gpu/*.rs`: `GPU` (gpu.rs:12-63),
`EpochKernel` (gpu/epoch.rs:10-125), `Results::to_gpu` / `update_from_gpu`
(scenario/results.rs:121-145) and `Measurements::to_gpu`
(data/shapes.rs:451-463). Every call goes to libcardiotrust_hip.so; there is
no CPU fallback: a missing library or device raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional, Sequence

import numpy as np

from .config import Algorithm, CAlgorithm

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcardiotrust_hip.so")

# Every symbol declared in include/cardiotrust_hip.h.
EXPORTS = (
    "ct_last_error", "ct_gpu_new", "ct_gpu_free", "ct_gpu_synchronize",
    "ct_measurements_to_gpu", "ct_buffer_free", "ct_results_to_gpu", "ct_results_gpu_free",
    "ct_results_update_from_gpu", "ct_results_simulate", "ct_results_simulate_noisy",
    "ct_measurement_matrix", "ct_model_connections", "ct_results_spherical",
    "ct_results_calculate_final_metrics", "ct_results_calculate_final_metrics_noalloc",
    "ct_epoch_kernel_last_loss", "ct_results_set_wait_limit",
    "ct_results_set_propagation", "ct_results_set_derivation",
    "ct_results_forward", "ct_results_backward", "ct_epoch_kernel_set_gradient",
    "ct_epoch_kernel_new",
    "ct_epoch_kernel_free", "ct_epoch_kernel_execute", "ct_epoch_kernel_set_freeze_delays",
    "ct_epoch_kernel_set_freeze_gains", "ct_epoch_kernel_set_beat_order",
    "ct_epoch_kernel_set_batch_index", "ct_epoch_kernel_get_batch_index",
    "ct_epoch_kernel_set_profiling", "ct_epoch_kernel_get_profile",
    "ct_results_objective_f64", "ct_epoch_kernel_set_share_propagation", "ct_batch_new", "ct_batch_free", "ct_batch_execute", "ct_batch_set_voxels_per_group",
    "ct_batch_set_graph", "ct_batch_set_profiling", "ct_batch_get_info",
)


class CTError(RuntimeError):
    pass


class _Desc(ctypes.Structure):
    _fields_ = [
        ("num_voxels", ctypes.c_int32), ("num_sensors", ctypes.c_int32),
        ("num_steps", ctypes.c_int32), ("num_beats", ctypes.c_int32),
        ("num_batch_entries", ctypes.c_int32), ("optimizer", ctypes.c_int32),
    ] + [(name, ctypes.c_void_p) for name in (
        "gains", "output_state_indices", "coefs", "delays", "initial_delays",
        "control_matrix", "control_function", "measurement_matrix")]


_HOST_FIELDS = (
    "gains", "coefs", "delays", "ap_outputs_now", "ap_outputs_last", "system_states",
    "measurements", "residuals", "d_gains", "d_coefs", "d_coefs_iir", "d_coefs_fir",
    "mapped_residuals", "maximum_regularization", "maximum_regularization_sum",
    "average_delays", "loss", "loss_mse", "loss_maximum_regularization", "loss_batch",
    "loss_mse_batch", "loss_maximum_regularization_batch", "derivatives_step",
)


class _Host(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in _HOST_FIELDS]


class Profile(ctypes.Structure):
    _fields_ = [
        ("propagate_ms", ctypes.c_double), ("propagate_launches", ctypes.c_int64),
        ("propagate_steps", ctypes.c_int64),
        ("project_ms", ctypes.c_double), ("regularize_ms", ctypes.c_double),
        ("transpose_ms", ctypes.c_double), ("map_residuals_ms", ctypes.c_double),
        ("derive_ms", ctypes.c_double), ("update_ms", ctypes.c_double),
        ("total_ms", ctypes.c_double), ("adjoint_sweep_ms", ctypes.c_double),
    ]

    def as_dict(self) -> Dict[str, float]:
        return {name: getattr(self, name) for name, _ in self._fields_}


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("scenarios", ctypes.c_int32), ("rounds", ctypes.c_int32), ("slots", ctypes.c_int32),
        ("voxels_per_group", ctypes.c_int32), ("jobs", ctypes.c_int32), ("graph", ctypes.c_int32),
        ("predicted_propagate_ms", ctypes.c_double), ("epoch_ms", ctypes.c_double),
        ("propagate_ms", ctypes.c_double), ("propagate_launches", ctypes.c_int64),
        ("voxel_steps", ctypes.c_int64), ("tuned_candidates", ctypes.c_int32),
        ("tuned_voxels", ctypes.c_int32 * 16), ("tuned_ms", ctypes.c_double * 16),
        ("side_streams", ctypes.c_int32), ("sweep_jobs", ctypes.c_int32),
    ]

    def as_dict(self) -> Dict[str, object]:
        d = {name: getattr(self, name) for name, _ in self._fields_[:11]}
        d["side_streams"] = self.side_streams
        d["sweep_jobs"] = self.sweep_jobs
        n = self.tuned_candidates
        d["autotune"] = {int(self.tuned_voxels[i]): float(self.tuned_ms[i]) for i in range(n)}
        return d


_lib = None


def lib() -> ctypes.CDLL:
    """Load libcardiotrust_hip.so (built by cardiotrust_amd/build.py)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CTError(f"{LIB_PATH} is not built; run cardiotrust_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, P = ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER
    sig = {
        "ct_last_error": ([], ctypes.c_char_p),
        "ct_gpu_new": ([i32, P(vp)], ctypes.c_int),
        "ct_gpu_free": ([vp], None),
        "ct_gpu_synchronize": ([vp], ctypes.c_int),
        "ct_measurements_to_gpu": ([vp, vp, i32, i32, i32, P(vp)], ctypes.c_int),
        "ct_buffer_free": ([vp], None),
        "ct_results_to_gpu": ([vp, P(_Desc), P(vp)], ctypes.c_int),
        "ct_results_gpu_free": ([vp], None),
        "ct_results_update_from_gpu": ([vp, P(_Host)], ctypes.c_int),
        "ct_results_simulate": ([vp, vp], ctypes.c_int),
        "ct_results_simulate_noisy": ([vp, vp, ctypes.c_uint64, vp], ctypes.c_int),
        "ct_results_spherical": ([vp, ctypes.c_float] + [vp] * 7, ctypes.c_int),
        "ct_results_calculate_final_metrics": ([vp, vp, vp, ctypes.c_int64] + [vp] * 4,
                                               ctypes.c_int),
        "ct_results_calculate_final_metrics_noalloc": ([vp, vp, vp, ctypes.c_int64] + [vp] * 4,
                                                       ctypes.c_int),
        "ct_epoch_kernel_last_loss": ([vp, vp, vp, vp, vp], ctypes.c_int),
        "ct_results_set_wait_limit": ([vp, ctypes.c_double], ctypes.c_int),
        "ct_measurement_matrix": ([vp, ctypes.c_int32, vp, ctypes.c_int32, vp, vp, ctypes.c_int32,
                                   vp, ctypes.c_float, vp], ctypes.c_int),
        "ct_model_connections": ([vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp, vp, vp,
                                  ctypes.c_float, ctypes.c_int32, vp, vp], ctypes.c_int),
        "ct_results_set_propagation": ([vp, i32, P(i32)], ctypes.c_int),
        "ct_results_set_derivation": ([vp, i32, P(i32)], ctypes.c_int),
        "ct_results_forward": ([vp, vp, i32, P(CAlgorithm), i32, P(ctypes.c_double)],
                               ctypes.c_int),
        "ct_results_backward": ([vp, vp, i32, P(CAlgorithm), i32, i32], ctypes.c_int),
        "ct_epoch_kernel_set_gradient": ([vp, i32, i32], ctypes.c_int),
        "ct_epoch_kernel_new": ([vp, vp, vp, P(CAlgorithm), i32, i32, i32, P(vp)], ctypes.c_int),
        "ct_epoch_kernel_free": ([vp], None),
        "ct_epoch_kernel_execute": ([vp], ctypes.c_int),
        "ct_epoch_kernel_set_freeze_delays": ([vp, i32], ctypes.c_int),
        "ct_epoch_kernel_set_freeze_gains": ([vp, i32], ctypes.c_int),
        "ct_epoch_kernel_set_beat_order": ([vp, vp, i32], ctypes.c_int),
        "ct_epoch_kernel_set_batch_index": ([vp, i32], ctypes.c_int),
        "ct_epoch_kernel_get_batch_index": ([vp, P(i32)], ctypes.c_int),
        "ct_epoch_kernel_set_profiling": ([vp, i32], ctypes.c_int),
        "ct_epoch_kernel_get_profile": ([vp, P(Profile)], ctypes.c_int),
        "ct_results_objective_f64": ([vp, vp, i32, P(CAlgorithm), i32, ctypes.c_int64,
                                      ctypes.c_double, P(ctypes.c_double)], ctypes.c_int),
        "ct_epoch_kernel_set_share_propagation": ([vp, i32], ctypes.c_int),
        "ct_batch_new": ([vp, P(vp), i32, P(vp)], ctypes.c_int),
        "ct_batch_free": ([vp], None),
        "ct_batch_execute": ([vp], ctypes.c_int),
        "ct_batch_set_voxels_per_group": ([vp, i32], ctypes.c_int),
        "ct_batch_set_graph": ([vp, i32], ctypes.c_int),
        "ct_batch_set_profiling": ([vp, i32], ctypes.c_int),
        "ct_batch_get_info": ([vp, P(BatchInfo)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().ct_last_error().decode(errors="replace")
        raise CTError(f"{what} failed ({rc}): {msg}")


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class GPU:
    """`GPU::new` (gpu.rs:31-63): one device and the library's stream."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(lib().ct_gpu_new(int(device), ctypes.byref(h)), "GPU::new")
        self.handle = h

    @classmethod
    def new(cls, device: int = 0) -> "GPU":
        return cls(device)

    def synchronize(self) -> None:
        _check(lib().ct_gpu_synchronize(self.handle), "synchronize")

    def close(self) -> None:
        if self.handle:
            lib().ct_gpu_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Buffer:
    """Device copy of measurements (`Measurements::to_gpu`, shapes.rs:451-463)."""

    def __init__(self, gpu: GPU, measurements: np.ndarray):
        m = np.ascontiguousarray(measurements, dtype=np.float32)
        if m.ndim != 3:
            raise ValueError("measurements must be (beats, steps, sensors)")
        h = ctypes.c_void_p()
        _check(lib().ct_measurements_to_gpu(gpu.handle, _ptr(m), *map(int, m.shape),
                                            ctypes.byref(h)), "Measurements::to_gpu")
        self.gpu, self.handle, self.shape = gpu, h, m.shape

    def close(self) -> None:
        if self.handle:
            lib().ct_buffer_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ResultsGPU:
    """Device state of one scenario (`ResultsGPU`, scenario/results.rs:36-41)."""

    def __init__(self, gpu: GPU, problem, num_batch_entries: int = 1, optimizer: int = 0):
        self.gpu = gpu
        self.N, self.S = problem.num_voxels, problem.num_sensors
        self.T, self.B = problem.num_steps, problem.num_beats
        self.nbatch = int(num_batch_entries)
        keep = lambda a, dt: np.ascontiguousarray(a, dtype=dt)
        self._arrays = dict(
            gains=keep(problem.gains, np.float32),
            output_state_indices=keep(problem.output_state_indices, np.int32),
            coefs=keep(problem.coefs, np.float32),
            delays=keep(problem.delays, np.int32),
            initial_delays=keep(problem.initial_delays, np.float32),
            control_matrix=keep(problem.control_matrix, np.float32),
            control_function=keep(problem.control_function, np.float32),
            measurement_matrix=keep(problem.measurement_matrix, np.float32),
        )
        d = _Desc(self.N, self.S, self.T, self.B, self.nbatch, int(optimizer),
                  *[_ptr(self._arrays[k]) for k in (
                      "gains", "output_state_indices", "coefs", "delays", "initial_delays",
                      "control_matrix", "control_function", "measurement_matrix")])
        h = ctypes.c_void_p()
        _check(lib().ct_results_to_gpu(gpu.handle, ctypes.byref(d), ctypes.byref(h)),
               "Results::to_gpu")
        self.handle = h
        self._arrays = None

    def update_from_gpu(self, fields: Optional[Sequence[str]] = None) -> Dict[str, np.ndarray]:
        """`Results::update_from_gpu`: download the requested arrays."""
        N, n, T, S, B = self.N, 3 * self.N, self.T, self.S, self.B
        shapes = dict(
            gains=((n, 78), np.float32), coefs=((N, 26), np.float32), delays=((N, 26), np.int32),
            ap_outputs_now=((n, 78), np.float32), ap_outputs_last=((n, 78), np.float32),
            system_states=((T, n), np.float32), measurements=((B, T, S), np.float32),
            residuals=((S,), np.float32), d_gains=((n, 78), np.float32),
            d_coefs=((N, 26), np.float32), d_coefs_iir=((n, 78), np.float32),
            d_coefs_fir=((n, 78), np.float32), mapped_residuals=((n,), np.float32),
            maximum_regularization=((n,), np.float32),
            maximum_regularization_sum=((1,), np.float32), average_delays=((N,), np.float32),
            loss=((T,), np.float32), loss_mse=((T,), np.float32),
            loss_maximum_regularization=((T,), np.float32),
            loss_batch=((self.nbatch,), np.float32), loss_mse_batch=((self.nbatch,), np.float32),
            loss_maximum_regularization_batch=((self.nbatch,), np.float32),
            derivatives_step=((1,), np.int64),
        )
        want = list(shapes) if fields is None else list(fields)
        out = {k: np.zeros(*shapes[k]) for k in want}
        h = _Host(*[_ptr(out.get(k)) for k in _HOST_FIELDS])
        _check(lib().ct_results_update_from_gpu(self.handle, ctypes.byref(h)),
               "Results::update_from_gpu")
        return out

    def set_propagation(self, mode: int) -> bool:
        """0 automatic, 1 one launch per step, 2 resident, 3 resident with the
        state rows exchanged through agent-scope fences; returns whether the
        resident kernel will run."""
        res = ctypes.c_int32()
        _check(lib().ct_results_set_propagation(self.handle, int(mode), ctypes.byref(res)),
               "set_propagation")
        return bool(res.value)

    def set_derivation(self, mode: int) -> bool:
        """0 automatic, 1 a lane per connection, 2 over voxel tiles; returns
        whether the tiled kernel will run."""
        res = ctypes.c_int32()
        _check(lib().ct_results_set_derivation(self.handle, int(mode), ctypes.byref(res)),
               "set_derivation")
        return bool(res.value)

    def objective_f64(self, actual: "Buffer", beat: int, config: Algorithm, param: int = 0,
                      index: int = 0, delta: float = 0.0) -> float:
        """FP64 objective with parameter `index` of `param` (1 gains, 2
        coefs, 0 none) shifted by `delta` (gradient checking)."""
        J = ctypes.c_double()
        c = config.to_c()
        _check(lib().ct_results_objective_f64(self.handle, actual.handle, int(beat),
                                              ctypes.byref(c), int(param), int(index),
                                              float(delta), ctypes.byref(J)), "objective_f64")
        return J.value

    def forward(self, actual: "Buffer", beat: int, config: Algorithm,
                checkpoint_interval: int = 64) -> float:
        """Forward pass of one beat (filter outputs stored every
        `checkpoint_interval` steps); returns the objective J."""
        J = ctypes.c_double()
        c = config.to_c()
        _check(lib().ct_results_forward(self.handle, actual.handle, int(beat), ctypes.byref(c),
                                        int(checkpoint_interval), ctypes.byref(J)), "forward")
        return J.value

    def backward(self, actual: "Buffer", beat: int, config: Algorithm, truncate: bool = False,
                 checkpoint_interval: int = 64) -> Dict[str, np.ndarray]:
        """Gradient of the last forward's J (or the reference's truncated
        derivation); returns d_gains (3N, 78) and d_coefs (N, 26)."""
        c = config.to_c()
        _check(lib().ct_results_backward(self.handle, actual.handle, int(beat), ctypes.byref(c),
                                         int(bool(truncate)), int(checkpoint_interval)),
               "backward")
        return self.update_from_gpu(["d_gains", "d_coefs"])

    def simulate(self, noise_std=None, seed: int = 0) -> np.ndarray:
        """Forward simulation of every beat; returns (B, T, S). With
        `noise_std` (S standard deviations) the measurements also get the
        library's seeded Gaussian noise (ct_results_simulate_noisy)."""
        out = np.zeros((self.B, self.T, self.S), dtype=np.float32)
        if noise_std is None:
            _check(lib().ct_results_simulate(self.handle, _ptr(out)), "simulate")
        else:
            sd = np.ascontiguousarray(noise_std, dtype=np.float32)
            if sd.shape != (self.S,):
                raise ValueError(f"noise_std must have shape ({self.S},)")
            _check(lib().ct_results_simulate_noisy(self.handle, _ptr(sd), ctypes.c_uint64(seed),
                                                   _ptr(out)), "simulate_noisy")
        return out

    def set_wait_limit(self, seconds: float) -> None:
        """Bound on every resident wait (ct_results_set_wait_limit)."""
        _check(lib().ct_results_set_wait_limit(self.handle, ctypes.c_double(seconds)),
               "set_wait_limit")

    def spherical(self, sample_rate_hz: float, full: bool = False) -> dict:
        """Spherical states of the device states (ct_results_spherical): the
        values at each voxel's maximum magnitude and the activation time in ms
        (N each), and with `full` the (T, N) magnitude, theta and phi."""
        N = self.N
        out = {k: np.zeros(N, np.float32) for k in ("magnitude_max", "theta_max", "phi_max",
                                                    "activation_time_ms")}
        if full:
            out.update({k: np.zeros((self.T, N), np.float32) for k in ("magnitude", "theta", "phi")})
        _check(lib().ct_results_spherical(
            self.handle, ctypes.c_float(sample_rate_hz), _ptr(out.get("magnitude")),
            _ptr(out.get("theta")), _ptr(out.get("phi")), _ptr(out["magnitude_max"]),
            _ptr(out["theta_max"]), _ptr(out["phi_max"]), _ptr(out["activation_time_ms"])),
            "spherical")
        return out

    FINAL_METRICS = ("dice_score_over_threshold", "iou_over_threshold",
                     "precision_over_threshold", "recall_over_threshold")

    def calculate_final(self, voxel_types, voxel_numbers, no_alloc: bool = False) -> dict:
        """metrics::calculate_final (algorithm/metrics.rs:225-241) over the
        device states (ct_results_calculate_final_metrics): the four (101,)
        arrays named as the reference's Metrics fields. voxel_types: VoxelType
        ordinals per grid cell (Pathological = 6); voxel_numbers: each cell's
        first state index, -1 for None (the reference's .npy forms).
        no_alloc: use ct_results_calculate_final_metrics_noalloc (the
        library's fixed device scratch instead of an allocation)."""
        types = np.ascontiguousarray(voxel_types, dtype=np.uint32).ravel()
        numbers = np.ascontiguousarray(voxel_numbers, dtype=np.int32).ravel()
        if types.shape != numbers.shape:
            raise CTError("voxel_types and voxel_numbers differ in shape")
        out = {k: np.zeros(101, np.float32) for k in self.FINAL_METRICS}
        fn = (lib().ct_results_calculate_final_metrics_noalloc if no_alloc
              else lib().ct_results_calculate_final_metrics)
        _check(fn(self.handle, _ptr(types), _ptr(numbers), types.size,
                  *(_ptr(out[k]) for k in self.FINAL_METRICS)), "calculate_final")
        return out

    def close(self) -> None:
        if self.handle:
            lib().ct_results_gpu_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class EpochKernel:
    """`EpochKernel` (gpu/epoch.rs:10-125) with `run_epoch` semantics."""

    def __init__(self, gpu: GPU, results: ResultsGPU, actual_measurements: Buffer,
                 config: Algorithm, number_of_states: int, number_of_sensors: int,
                 number_of_steps: int):
        h = ctypes.c_void_p()
        c = config.to_c()
        _check(lib().ct_epoch_kernel_new(gpu.handle, results.handle, actual_measurements.handle,
                                         ctypes.byref(c), int(number_of_states),
                                         int(number_of_sensors), int(number_of_steps),
                                         ctypes.byref(h)), "EpochKernel::new")
        self.gpu, self.results, self.actual, self.handle = gpu, results, actual_measurements, h

    def execute(self) -> None:
        _check(lib().ct_epoch_kernel_execute(self.handle), "EpochKernel::execute")

    def set_gradient(self, mode: int, checkpoint_interval: int = 64) -> None:
        """0 = the reference's derivation, 1 = exact reverse-mode gradient."""
        _check(lib().ct_epoch_kernel_set_gradient(self.handle, int(mode), int(checkpoint_interval)),
               "set_gradient")

    def set_share_propagation(self, enabled: Optional[bool]) -> None:
        """Propagate once per update batch (the beats' states are identical);
        with the reverse-mode gradient, one backward pass per update batch.
        None = automatic (the default: shared under both gradients; the
        results are bit-identical under the reference derivation and the same
        gradient up to FP32 rounding under the reverse-mode one)."""
        mode = -1 if enabled is None else int(bool(enabled))
        _check(lib().ct_epoch_kernel_set_share_propagation(self.handle, mode),
               "set_share_propagation")

    def set_freeze_delays(self, value: bool) -> None:
        _check(lib().ct_epoch_kernel_set_freeze_delays(self.handle, int(bool(value))),
               "set_freeze_delays")

    def set_freeze_gains(self, value: bool) -> None:
        _check(lib().ct_epoch_kernel_set_freeze_gains(self.handle, int(bool(value))),
               "set_freeze_gains")

    def set_beat_order(self, order: Optional[Sequence[int]]) -> None:
        o = None if order is None else np.ascontiguousarray(order, dtype=np.int32)
        _check(lib().ct_epoch_kernel_set_beat_order(self.handle, _ptr(o),
                                                    0 if o is None else len(o)),
               "set_beat_order")

    def set_batch_index(self, index: int) -> None:
        _check(lib().ct_epoch_kernel_set_batch_index(self.handle, int(index)), "set_batch_index")

    def batch_index(self) -> int:
        v = ctypes.c_int32()
        _check(lib().ct_epoch_kernel_get_batch_index(self.handle, ctypes.byref(v)),
               "get_batch_index")
        return int(v.value)

    def last_loss(self) -> dict:
        """The last recorded batch's losses and whether the loss is a normal
        number (the reference's divergence stop; ct_epoch_kernel_last_loss)."""
        v = [ctypes.c_float() for _ in range(3)]
        ok = ctypes.c_int32()
        _check(lib().ct_epoch_kernel_last_loss(self.handle, *map(ctypes.byref, v),
                                               ctypes.byref(ok)), "last_loss")
        return {"loss": v[0].value, "loss_mse": v[1].value,
                "loss_maximum_regularization": v[2].value, "is_normal": bool(ok.value)}

    def set_profiling(self, enabled: bool) -> None:
        _check(lib().ct_epoch_kernel_set_profiling(self.handle, int(bool(enabled))),
               "set_profiling")

    def profile(self) -> Dict[str, float]:
        p = Profile()
        _check(lib().ct_epoch_kernel_get_profile(self.handle, ctypes.byref(p)), "get_profile")
        return p.as_dict()

    def close(self) -> None:
        if self.handle:
            lib().ct_epoch_kernel_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def measurement_matrix(gpu: GPU, voxel_positions_mm, sensor_positions_mm, sensor_orientations,
                       sensor_offsets_mm, voxel_size_mm: float) -> np.ndarray:
    """(B, S, 3N) Biot-Savart matrix built on the GPU (ct_measurement_matrix)."""
    vp = np.ascontiguousarray(voxel_positions_mm, dtype=np.float32).reshape(-1, 3)
    sp = np.ascontiguousarray(sensor_positions_mm, dtype=np.float32).reshape(-1, 3)
    so = np.ascontiguousarray(sensor_orientations, dtype=np.float32).reshape(-1, 3)
    off = np.ascontiguousarray(sensor_offsets_mm, dtype=np.float32).reshape(-1, 3)
    out = np.zeros((off.shape[0], sp.shape[0], 3 * vp.shape[0]), dtype=np.float32)
    _check(lib().ct_measurement_matrix(gpu.handle, vp.shape[0], _ptr(vp), sp.shape[0], _ptr(sp),
                                       _ptr(so), off.shape[0], _ptr(off),
                                       ctypes.c_float(voxel_size_mm), _ptr(out)),
           "measurement_matrix")
    return out


def model_connections(gpu: GPU, state_numbers, positions_mm, velocities_m_per_s,
                      sample_rate_hz: float, num_voxels: int):
    """(output_state_indices (3N, 78) int32, delay_samples (N, 26) float32) of a
    voxel grid, built on the GPU (ct_model_connections)."""
    num = np.ascontiguousarray(state_numbers, dtype=np.int32)
    nx, ny, nz = num.shape
    pos = np.ascontiguousarray(positions_mm, dtype=np.float32).reshape(nx, ny, nz, 3)
    vel = np.ascontiguousarray(velocities_m_per_s, dtype=np.float32).reshape(nx, ny, nz)
    osi = np.empty((3 * num_voxels, 78), dtype=np.int32)
    dly = np.empty((num_voxels, 26), dtype=np.float32)
    _check(lib().ct_model_connections(gpu.handle, nx, ny, nz, _ptr(num), _ptr(pos), _ptr(vel),
                                      ctypes.c_float(sample_rate_hz), num_voxels, _ptr(osi),
                                      _ptr(dly)), "model_connections")
    return osi, dly


def simulate(gpu: GPU, problem) -> np.ndarray:
    """Noise-free measurements of `problem`'s parameters, computed on the GPU."""
    r = ResultsGPU(gpu, problem)
    try:
        return r.simulate()
    finally:
        r.close()


class Batch:
    """Many scenarios advanced one epoch per `execute` (ct_batch_*): each
    member's results equal those of its own EpochKernel.execute()."""

    def __init__(self, gpu: GPU, kernels: Sequence["EpochKernel"], voxels_per_group: int = 0,
                 graph: bool = False):
        arr = (ctypes.c_void_p * len(kernels))(*[k.handle for k in kernels])
        h = ctypes.c_void_p()
        _check(lib().ct_batch_new(gpu.handle, arr, len(kernels), ctypes.byref(h)), "ct_batch_new")
        self.handle = h
        self.kernels = list(kernels)  # keep the members alive
        self.set_voxels_per_group(voxels_per_group)
        self.set_graph(graph)

    def execute(self) -> None:
        _check(lib().ct_batch_execute(self.handle), "ct_batch_execute")

    def set_voxels_per_group(self, voxels: int) -> None:
        _check(lib().ct_batch_set_voxels_per_group(self.handle, int(voxels)),
               "set_voxels_per_group")

    def set_graph(self, enabled: bool) -> None:
        _check(lib().ct_batch_set_graph(self.handle, int(bool(enabled))), "set_graph")

    def set_profiling(self, enabled: bool) -> None:
        _check(lib().ct_batch_set_profiling(self.handle, int(bool(enabled))), "set_profiling")

    def info(self) -> Dict[str, object]:
        out = BatchInfo()
        _check(lib().ct_batch_get_info(self.handle, ctypes.byref(out)), "get_info")
        return out.as_dict()

    def close(self) -> None:
        if self.handle:
            lib().ct_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
