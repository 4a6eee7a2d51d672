"""Seeded synthetic CardioTrust problems (geometry, parameters, sensors, drive).

The reference builds its model from a voxelised heart (`model/spatial/*`,
`model/functional/allpass.rs:71-107,212-289`, `measurement.rs:66-123`,
`control.rs:43-64`). The epoch workloads measured here use synthetic cubes and
ellipsoidal shells instead, built with the same array layouts:

* voxels are numbered in C order over (x, y, z) of the active mask, so voxel
  `v` owns states 3v, 3v+1, 3v+2 (`voxels.rs:407`);
* connection `d` of voxel `v` points to the neighbour at the offset
  `delay_index_to_offset(d)` (`allpass.rs:603-618`); `output_state_indices`
  holds 3w+out for every (3v+in, 3d+out) and -1 where the neighbour is missing
  (`allpass.rs:212-289`);
* the measurement matrix is Biot-Savart from voxel centres to sensor positions
  (`measurement.rs:85-117`), evaluated in float64 and rounded once to float32;
* the control matrix has 1.0 at the x-state of the driven voxel
  (`control.rs:56-60`).

All randomness comes from splitmix64 over (seed, stream, counter), so every
array is reproducible from the seed alone on any machine. Normal deviates use
Box-Muller on two uniforms.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

# physical_constants 0.5.0 VACUUM_MAG_PERMEABILITY (CODATA 2018), N/A^2.
MU0 = 1.25663706212e-6

# (dx, dy, dz) of connection d, allpass.rs:603-618 (delay_index_to_offset).
OFFSETS = np.array(
    [
        [(c // 9) % 3 - 1, (c // 3) % 3 - 1, c % 3 - 1]
        for c in [d + 1 if d > 12 else d for d in range(26)]
    ],
    dtype=np.int64,
)

def splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser applied to uint64 counters (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        t = z >> np.uint64(30)
        z ^= t
        z *= np.uint64(0xBF58476D1CE4E5B9)
        np.right_shift(z, np.uint64(27), out=t)
        z ^= t
        z *= np.uint64(0x94D049BB133111EB)
        np.right_shift(z, np.uint64(31), out=t)
        z ^= t
        return z

def uniform(seed: int, stream: int, count: int, offset: int = 0) -> np.ndarray:
    """`count` float64 uniforms in [0, 1) for (seed, stream)."""
    key = splitmix64(np.array([(seed << 20) ^ stream], dtype=np.uint64))[0]
    ctr = np.arange(offset, offset + count, dtype=np.uint64)
    with np.errstate(over="ignore"):
        bits = splitmix64(ctr ^ key)
    return (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

def normal(seed: int, stream: int, count: int) -> np.ndarray:
    """`count` standard normal float64 deviates (Box-Muller)."""
    u = uniform(seed, stream, 2 * count)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)

def measurement_noise(seed: int, sensor_std: np.ndarray, num_beats: int,
                      num_steps: int) -> np.ndarray:
    """(B, T, S) float32 noise of ct_results_simulate_noisy for `seed`: element
    (b, t, s) is float32(sensor_std[s] * z_k), k = (s B + b) T + t (the
    reference's sensor, beat, step order, data/simulation.rs:152-169), z_k
    the Box-Muller deviate of counter uniforms 2k and 2k+1 of stream 1000 of
    `seed` (key taken modulo 2^64)."""
    sd = np.asarray(sensor_std, dtype=np.float32)
    S, B, T = sd.shape[0], int(num_beats), int(num_steps)
    key = splitmix64(np.array([((int(seed) << 20) ^ 1000) & 0xFFFFFFFFFFFFFFFF],
                              dtype=np.uint64))[0]
    ctr = np.arange(0, 2 * S * B * T, dtype=np.uint64)
    with np.errstate(over="ignore"):
        bits = splitmix64(ctr ^ key)
    u = (bits >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    z = np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(2.0 * np.pi * u[1::2])
    noise = (sd.astype(np.float64)[:, None, None] * z.reshape(S, B, T)).astype(np.float32)
    return np.ascontiguousarray(noise.transpose(1, 2, 0))

@dataclass
class Geometry:
    """Active voxels of a regular grid, numbered in C order."""

    shape: tuple
    voxel_size_mm: float
    mask: np.ndarray  # (nx, ny, nz) bool
    numbers: np.ndarray  # (nx, ny, nz) int64, -1 where inactive
    coords: np.ndarray  # (N, 3) int64
    positions_mm: np.ndarray  # (N, 3) float64, voxel centres
    neighbours: np.ndarray  # (N, 26) int32, -1 where missing

    @property
    def num_voxels(self) -> int:
        return int(self.coords.shape[0])

def grid_geometry(mask: np.ndarray, voxel_size_mm: float) -> Geometry:
    mask = np.asarray(mask, dtype=bool)
    shape = mask.shape
    numbers = np.full(shape, -1, dtype=np.int64)
    coords = np.argwhere(mask)  # C order
    numbers[tuple(coords.T)] = np.arange(coords.shape[0])
    neighbours = np.full((coords.shape[0], 26), -1, dtype=np.int32)
    for d in range(26):
        c = coords + OFFSETS[d]
        ok = np.all((c >= 0) & (c < np.array(shape)), axis=1)
        w = np.full(coords.shape[0], -1, dtype=np.int64)
        w[ok] = numbers[tuple(c[ok].T)]
        neighbours[:, d] = w
    positions = (coords.astype(np.float64) + 0.5) * voxel_size_mm
    return Geometry(shape, voxel_size_mm, mask, numbers, coords, positions, neighbours)

def cube(n: int, voxel_size_mm: float = 2.5) -> Geometry:
    return grid_geometry(np.ones((n, n, n), dtype=bool), voxel_size_mm)

def ellipsoid_shell(shape=(64, 64, 64), outer=(32.0, 32.0, 32.0), inner=(16.0, 16.0, 16.0),
                    voxel_size_mm: float = 2.5) -> Geometry:
    """Voxels whose centre lies inside the outer and outside the inner ellipsoid
    (semi-axes in voxels, both centred on the grid centre)."""
    idx = np.indices(shape).astype(np.float64) + 0.5
    centre = np.array(shape, dtype=np.float64)[:, None, None, None] / 2.0
    rel = idx - centre
    ro = sum((rel[k] / outer[k]) ** 2 for k in range(3))
    ri = sum((rel[k] / inner[k]) ** 2 for k in range(3))
    return grid_geometry((ro <= 1.0) & (ri > 1.0), voxel_size_mm)

def blob(seed: int, num_voxels: int, voxel_size_mm: float = 2.5) -> Geometry:
    """Seeded random blob of exactly `num_voxels` voxels: the `num_voxels`
    highest cells of a sum of six Gaussian bumps (random centres and widths)
    in a cubic box of about 1.8 x `num_voxels` cells."""
    side = max(2, int(np.ceil((1.8 * num_voxels) ** (1.0 / 3.0))))
    u = uniform(seed, 20, 24).reshape(6, 4)
    idx = np.indices((side, side, side)).reshape(3, -1).T.astype(np.float64) + 0.5
    field = np.zeros(idx.shape[0])
    for c in u:
        centre = (0.2 + 0.6 * c[:3]) * side
        width = (0.15 + 0.2 * c[3]) * side
        field += np.exp(-np.sum((idx - centre) ** 2, axis=1) / (2.0 * width * width))
    # ties broken by cell number, so the count is exact
    keep = np.lexsort((np.arange(field.size), -field))[:num_voxels]
    mask = np.zeros(field.size, dtype=bool)
    mask[keep] = True
    return grid_geometry(mask.reshape(side, side, side), voxel_size_mm)

def sphere_region(geometry: Geometry, seed: int, radius_fraction: float = 0.2) -> np.ndarray:
    """(N,) bool: voxels within a seeded sphere (radius `radius_fraction` of
    the grid's smallest side) centred on a random active voxel."""
    if geometry.num_voxels == 0:
        return np.zeros(0, dtype=bool)
    pick = int(uniform(seed, 30, 1)[0] * geometry.num_voxels)
    centre = geometry.coords[pick].astype(np.float64)
    r = radius_fraction * min(geometry.shape)
    return np.sum((geometry.coords - centre) ** 2, axis=1) <= r * r

def grid_motion_offsets(steps=(1, 1, 1), range_mm=(0.0, 0.0, 0.0)) -> np.ndarray:
    """(B, 3) sensor-array offsets of SensorArrayMotion::Grid
    (`sensors.rs:194-225`): B = steps_x * steps_y * steps_z positions, x
    outermost, each axis spanning `range_mm` in steps - 1 equal steps."""
    step = [r / (n - 1) if n > 1 else 0.0 for r, n in zip(range_mm, steps)]
    out = [[step[0] * x, step[1] * y, step[2] * z]
           for x in range(steps[0]) for y in range(steps[1]) for z in range(steps[2])]
    return np.array(out, dtype=np.float32).astype(np.float64)

def output_state_indices(neighbours: np.ndarray) -> np.ndarray:
    """(3N, 78) int32: 3w+out at (3v+in, 3d+out), -1 where no neighbour."""
    N = neighbours.shape[0]
    out = np.arange(3, dtype=np.int64)
    base = neighbours.astype(np.int64)[:, :, None] * 3 + out[None, None, :]  # (N,26,3)
    base = np.where(neighbours[:, :, None] >= 0, base, -1).reshape(N, 78)
    return np.repeat(base, 3, axis=0).astype(np.int32)

@dataclass
class Sensors:
    positions_mm: np.ndarray  # (S, 3)
    orientations: np.ndarray  # (S, 3)
    beat_offsets_mm: np.ndarray  # (B, 3)

def sensor_grid(geometry: Geometry, per_axis: int = 6, pitch_mm: float = 30.0,
                height_mm: float = 100.0, num_beats: int = 1,
                beat_offsets_mm: Optional[np.ndarray] = None) -> Sensors:
    """`per_axis` x `per_axis` grid of 3-axis magnetometers centred over the
    volume, `height_mm` above its top face; channel order (x, y, axis) with
    axis = channel % 3 as `sensors.rs:68-87`."""
    extent = np.array(geometry.shape, dtype=np.float64) * geometry.voxel_size_mm
    cx, cy, top = extent[0] / 2.0, extent[1] / 2.0, extent[2]
    pos, ori = [], []
    for ix in range(per_axis):
        for iy in range(per_axis):
            p = [cx + (ix - (per_axis - 1) / 2.0) * pitch_mm,
                 cy + (iy - (per_axis - 1) / 2.0) * pitch_mm, top + height_mm]
            for axis in range(3):
                pos.append(p)
                o = [0.0, 0.0, 0.0]
                o[axis] = 1.0
                ori.append(o)
    if beat_offsets_mm is None:
        beat_offsets_mm = np.zeros((num_beats, 3))
    return Sensors(np.array(pos), np.array(ori), np.asarray(beat_offsets_mm, dtype=np.float64))

def measurement_matrix(geometry: Geometry, sensors: Sensors) -> np.ndarray:
    """(B, S, 3N) float32 Biot-Savart matrix, `measurement.rs:85-117`:
    H[b,s,3v+k] = c * (d x o)_k / |d|^3 with d = (sensor + offset_b - voxel)/1000
    and c = mu0 * V / (4 pi) * 1e12 (V in m^3, result in pT per A m^2)."""
    V = (geometry.voxel_size_mm / 1000.0) ** 3
    c = (MU0 * V) / (4.0 * np.pi) * 1e12
    B = sensors.beat_offsets_mm.shape[0]
    S = sensors.positions_mm.shape[0]
    N = geometry.num_voxels
    H = np.empty((B, S, 3 * N), dtype=np.float32)
    vp = geometry.positions_mm
    for b in range(B):
        for s in range(S):
            sp = sensors.positions_mm[s] + sensors.beat_offsets_mm[b]
            d = (sp[None, :] - vp) / 1000.0
            r3 = np.sqrt(np.sum(d * d, axis=1)) ** 3
            o = sensors.orientations[s]
            cross = np.stack([o[2] * d[:, 1] - o[1] * d[:, 2],
                              o[0] * d[:, 2] - o[2] * d[:, 0],
                              o[1] * d[:, 0] - o[0] * d[:, 1]], axis=1)
            H[b, s] = (c * cross / r3[:, None]).reshape(-1).astype(np.float32)
    return H

def gaussian_derivative_pulse(num_steps: int, sample_rate_hz: float = 2000.0,
                              t0_ms: float = 20.0, sigma_ms: float = 4.0) -> np.ndarray:
    """(T,) float32 derivative-of-Gaussian drive, peak-normalised to |u| = 1."""
    t_ms = np.arange(num_steps, dtype=np.float64) * (1000.0 / sample_rate_hz)
    g = -(t_ms - t0_ms) / sigma_ms**2 * np.exp(-((t_ms - t0_ms) ** 2) / (2.0 * sigma_ms**2))
    return (g / np.max(np.abs(g))).astype(np.float32)

def tiled_control_function(raw: np.ndarray, num_steps: int, sample_rate_hz: float = 2000.0,
                           raw_rate_hz: float = 2000.0) -> np.ndarray:
    """(T,) float32 drive from a recorded control function, as
    ControlFunction::from_model_config builds its Ohara drive
    (model/functional/control.rs:174-223) at the recording's own rate: sample i
    is raw[i % len(raw)]. The reference resamples other rates with a sinc
    resampler (rubato), which is not provided here: ValueError."""
    raw = np.asarray(raw, dtype=np.float32).ravel()
    if raw.size == 0:
        raise ValueError("empty control function")
    if abs(sample_rate_hz - raw_rate_hz) > 1e-3 * max(abs(sample_rate_hz), abs(raw_rate_hz)):
        raise ValueError("resampling a recorded control function is not provided")
    return raw[np.arange(num_steps) % raw.size].copy()

def coef_to_samples(coef: np.ndarray) -> np.ndarray:
    """allpass.rs:665-668 in float32."""
    coef = np.asarray(coef, dtype=np.float32)
    one = np.float32(1.0)
    return (one - coef) / (coef + one)

@dataclass
class Problem:
    """Reference-layout arrays of one scenario (see module docstring)."""

    geometry: Geometry
    num_sensors: int
    num_steps: int
    num_beats: int
    gains: np.ndarray  # (3N, 78) f32, initial estimate
    output_state_indices: np.ndarray  # (3N, 78) i32
    coefs: np.ndarray  # (N, 26) f32, initial estimate
    delays: np.ndarray  # (N, 26) i32
    initial_delays: np.ndarray  # (N, 26) f32
    control_matrix: np.ndarray  # (3N,) f32
    control_function: np.ndarray  # (T,) f32
    measurement_matrix: np.ndarray  # (B, S, 3N) f32
    measurements: Optional[np.ndarray] = None  # (B, T, S) f32, from the truth
    true_gains: Optional[np.ndarray] = None
    true_coefs: Optional[np.ndarray] = None
    true_delays: Optional[np.ndarray] = None
    meta: dict = field(default_factory=dict)

    @property
    def num_voxels(self) -> int:
        return self.geometry.num_voxels

    @property
    def num_states(self) -> int:
        return 3 * self.geometry.num_voxels

    def truth(self) -> "Problem":
        """The same problem with the ground-truth parameters in place."""
        return Problem(self.geometry, self.num_sensors, self.num_steps, self.num_beats,
                       self.true_gains, self.output_state_indices, self.true_coefs,
                       self.true_delays, self.initial_delays, self.control_matrix,
                       self.control_function, self.measurement_matrix, self.measurements,
                       self.true_gains, self.true_coefs, self.true_delays, dict(self.meta))

def make_problem(geometry: Geometry, num_steps: int, max_delay: int, seed: int,
                 num_beats: int = 1, gain_sigma: float = 0.1,
                 perturbation: float = 0.2,
                 beat_offsets_mm: Optional[np.ndarray] = None,
                 slow_region: Optional[np.ndarray] = None,
                 simulate: Optional[Callable[["Problem"], np.ndarray]] = None,
                 sensors_per_axis: int = 6) -> Problem:
    """Random ground-truth parameters, their perturbed initial estimate and,
    when `simulate` is given, the noise-free measurements of the truth.

    delays ~ U{1..max_delay}, coefs ~ U(0.05, 0.95) and gains ~ N(0, sigma^2)
    with sigma = `gain_sigma` (the "N(0, 0.1)" of the workload read as a
    standard deviation) on every existing connection; missing connections keep
    gain 0, delay 1 and coef 0.5. The estimate multiplies gains and coefs by
    (1 + perturbation * N(0, 1)), clamps coefs to [1e-4, 1 - 1e-4] and keeps
    the true integer delays. With `slow_region` ((N,) bool) the true delays of
    the connections of those voxels are doubled (a pathological, slowly
    conducting region) while the estimate starts from the undoubled ones.
    """
    N = geometry.num_voxels
    nb = geometry.neighbours
    exists = nb >= 0
    D = (1 + np.floor(uniform(seed, 1, N * 26) * max_delay)).astype(np.int32).reshape(N, 26)
    a = (0.05 + 0.9 * uniform(seed, 2, N * 26)).astype(np.float32).reshape(N, 26)
    D = np.where(exists, D, 1).astype(np.int32)
    D_estimate = D.copy()
    if slow_region is not None:
        D = np.where(exists & np.asarray(slow_region, dtype=bool)[:, None], 2 * D, D)
        D = D.astype(np.int32)
    a = np.where(exists, a, np.float32(0.5)).astype(np.float32)
    g = (gain_sigma * normal(seed, 3, N * 78 * 3)).astype(np.float32).reshape(3 * N, 78)
    conn = np.repeat(np.repeat(exists, 3, axis=1), 3, axis=0)  # (3N, 78)
    g = np.where(conn, g, np.float32(0.0)).astype(np.float32)

    eg = normal(seed, 4, N * 78 * 3).reshape(3 * N, 78)
    ea = normal(seed, 5, N * 26).reshape(N, 26)
    g0 = (g.astype(np.float64) * (1.0 + perturbation * eg)).astype(np.float32)
    a0 = np.clip(a.astype(np.float64) * (1.0 + perturbation * ea), 1e-4, 1.0 - 1e-4)
    a0 = np.where(exists, a0, 0.5).astype(np.float32)

    sensors = sensor_grid(geometry, per_axis=sensors_per_axis, num_beats=num_beats,
                          beat_offsets_mm=beat_offsets_mm)
    H = measurement_matrix(geometry, sensors)
    u = gaussian_derivative_pulse(num_steps)
    b = np.zeros(3 * N, dtype=np.float32)
    if N > 0:
        b[0] = 1.0
    p = Problem(
        geometry=geometry,
        num_sensors=int(H.shape[1]),
        num_steps=int(num_steps),
        num_beats=int(num_beats),
        gains=g0,
        output_state_indices=output_state_indices(nb),
        coefs=a0,
        delays=D_estimate.copy(),
        initial_delays=(D_estimate.astype(np.float32) + coef_to_samples(a0)).astype(np.float32),
        control_matrix=b,
        control_function=u,
        measurement_matrix=H,
        true_gains=g,
        true_coefs=a,
        true_delays=D,
        meta={"seed": seed, "max_delay": max_delay},
    )
    if simulate is not None:
        p.measurements = np.ascontiguousarray(simulate(p.truth()), dtype=np.float32)
    return p

def config0(simulate=None) -> Problem:
    """10^3 cube of 2.5 mm voxels, delays 1-20, T = 1000 at 2 kHz, one beat."""
    return make_problem(cube(10), num_steps=1000, max_delay=20, seed=0, simulate=simulate)

def config1(simulate=None, num_steps: int = 2000) -> Problem:
    """64^3 grid masked to an ellipsoidal shell (outer semi-axes 32, inner 16
    voxels; N = 120,000 active voxels), delays 1-40, T = 2000, one beat."""
    return make_problem(ellipsoid_shell(), num_steps=num_steps, max_delay=40, seed=1,
                        simulate=simulate)

def config2(simulate=None, num_steps: int = 2000) -> Problem:
    """32^3 full cube (N = 32,768), delays 1-40, T = 2000, one beat."""
    return make_problem(cube(32), num_steps=num_steps, max_delay=40, seed=2, simulate=simulate)

def _uneven_scenario(seed: int, i: int, n: int, beats: int, steps: int) -> Problem:
    g = blob(seed * 1000 + i, n)
    offsets = grid_motion_offsets((beats, 1, 1), (10.0 * (beats - 1), 0.0, 0.0))
    return make_problem(g, num_steps=steps, max_delay=20, seed=seed * 1000 + i,
                        num_beats=beats, beat_offsets_mm=offsets)

def _shell_scenario(seed: int, i: int, ux: float, uy: float, beats: int,
                    num_steps: int) -> Problem:
    outer = (32.0, 32.0 - 2.0 * ux, 32.0)
    inner = (16.0 + 2.0 * uy, 16.0, 16.0)
    g = ellipsoid_shell(outer=outer, inner=inner)
    region = sphere_region(g, seed * 1000 + i)
    offsets = grid_motion_offsets((beats, 1, 1), (10.0 * (beats - 1), 0.0, 0.0))
    return make_problem(g, num_steps=num_steps, max_delay=40, seed=seed * 1000 + i,
                        num_beats=beats, beat_offsets_mm=offsets, slow_region=region)

def _build(fn, specs, simulate, workers: int) -> list:
    """fn(*spec) for every spec, in `workers` spawned processes when > 1 (they
    run numpy only; the arrays are the same as built in this process), then
    the ground truth's measurements with `simulate` when given."""
    if workers > 1 and len(specs) > 1:
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(max_workers=min(workers, len(specs)),
                                 mp_context=mp.get_context("spawn")) as pool:
            out = list(pool.map(fn, *zip(*specs)))
    else:
        out = [fn(*spec) for spec in specs]
    if simulate is not None:
        for p in out:
            p.measurements = np.ascontiguousarray(simulate(p.truth()), dtype=np.float32)
    return out

def config3(simulate=None, count: int = 64, seed: int = 3, min_voxels: int = 1000,
            max_voxels: int = 40000, min_steps: int = 600, max_steps: int = 2000,
            max_beats: int = 8, workers: int = 1) -> list:
    """`count` scenarios of uneven size: voxels log-uniform in
    [min_voxels, max_voxels] (seeded blobs, 2.5 mm), beats uniform in
    1..max_beats (sensor array moved along x by 10 mm per beat, grid motion),
    steps uniform in [min_steps, max_steps] per scenario (the reference's
    Measurements are (B, T, S): one length per scenario), delays 1-20.
    `workers` > 1 builds the scenarios in that many processes."""
    u = uniform(seed, 40, 3 * count).reshape(count, 3)
    specs = []
    for i in range(count):
        n = int(round(np.exp(np.log(min_voxels) + u[i, 0] * np.log(max_voxels / min_voxels))))
        beats = 1 + int(u[i, 1] * max_beats)
        steps = min_steps + int(u[i, 2] * (max_steps - min_steps + 1))
        specs.append((seed, i, n, beats, steps))
    return _build(_uneven_scenario, specs, simulate, workers)

def config4(simulate=None, count: int = 8, seed: int = 4, beats: int = 4,
            num_steps: int = 2000, workers: int = 1) -> list:
    """`count` ellipsoidal shells of about 120k voxels (64^3 grid, semi-axes
    varied by a few voxels per scenario) with a seeded slow sub-region (true
    delays doubled), `beats` beats with grid sensor motion, T = num_steps,
    delays 1-40. `workers` > 1 builds the scenarios in that many processes."""
    u = uniform(seed, 50, 2 * count).reshape(count, 2)
    specs = [(seed, i, float(u[i, 0]), float(u[i, 1]), beats, num_steps) for i in range(count)]
    return _build(_shell_scenario, specs, simulate, workers)
