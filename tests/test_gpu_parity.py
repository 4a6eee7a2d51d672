"""Parity of the gfx950 epoch with the CPU oracle, through the C ABI.

Bars (written per assertion):
  * states, filter outputs, integer delays, neighbour-derived indices: bit-exact;
  * measurements, residual-derived losses and derivatives: relative error
    ||gpu - oracle||_inf / ||oracle||_inf <= 1e-4 (FP32; the reductions H·x and
    Hᵀr run on MFMA in a different summation order than ndarray/matrixmultiply);
  * every compared array also reports its error against the FP64 oracle.
"""
import os
import sys

import numpy as np
import pytest

from cardiotrust_amd import Algorithm, APDerivative, Optimizer, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

REL = 1e-4


def rel_inf(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    if scale == 0.0:
        return float(np.max(np.abs(a))) if a.size else 0.0
    return float(np.max(np.abs(a - b)) / scale)


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    from cardiotrust_amd.gpu import GPU
    g = GPU(0)
    yield g
    g.close()


def run_gpu_epochs(gpu, p, cfg, epochs=1, beat_order=None, nbatch=1, optimizer=0, profile=False,
                   mode=0):
    from cardiotrust_amd.gpu import Buffer, EpochKernel, ResultsGPU
    r = ResultsGPU(gpu, p, num_batch_entries=nbatch, optimizer=optimizer)
    resident = r.set_propagation(mode)
    if mode == 2:
        assert resident
    buf = Buffer(gpu, p.measurements)
    k = EpochKernel(gpu, r, buf, cfg, 3 * p.num_voxels, p.num_sensors, p.num_steps)
    if beat_order is not None:
        k.set_beat_order(beat_order)
    k.set_profiling(profile)
    for _ in range(epochs):
        k.execute()
    out = r.update_from_gpu()
    out["batch_index"] = k.batch_index()
    prof = k.profile() if profile else None
    k.close()
    buf.close()
    r.close()
    return out, prof


def run_oracle_epochs(p, cfg, dtype, epochs=1, beat_order=None, nbatch=1, optimizer=0):
    o = oracle.OracleResults(p, dtype, optimizer=optimizer, num_batch_entries=nbatch)
    for _ in range(epochs):
        o.run_epoch(cfg, beat_order)
    return o


EXACT = ("system_states", "ap_outputs_now", "ap_outputs_last", "delays")
CLOSE = ("measurements", "gains", "coefs", "d_gains", "d_coefs", "d_coefs_fir", "d_coefs_iir",
         "residuals", "mapped_residuals", "loss", "loss_mse", "loss_maximum_regularization",
         "loss_batch", "loss_mse_batch", "loss_maximum_regularization_batch",
         "maximum_regularization", "maximum_regularization_sum")


def compare(g, o32, o64, exact=EXACT, close=CLOSE, rel=REL):
    report = {}
    for k in exact:
        want = getattr(o32, k)
        assert np.array_equal(g[k].view(np.uint32) if g[k].dtype == np.float32 else g[k],
                              want.view(np.uint32) if want.dtype == np.float32 else want), k
    for k in close:
        want = getattr(o32, k)
        e32 = rel_inf(g[k], want)
        e64 = rel_inf(g[k], getattr(o64, k))
        o_e64 = rel_inf(want, getattr(o64, k))
        report[k] = (e32, e64, o_e64)
        assert e32 <= rel, (k, e32)
    note("  rel err vs oracle f32 / vs f64 (oracle f32 vs f64): " +
         ", ".join(f"{k}={v[0]:.1e}/{v[1]:.1e}({v[2]:.1e})" for k, v in report.items()))
    return report


@pytest.fixture(scope="module")
def config0():
    return synth.config0(simulate=oracle.simulate_f32)


@pytest.mark.parametrize("mode", [1, 2], ids=["per-step", "resident"])
def test_epoch_config0_parity(gpu, config0, mode):
    note(f"case: config0 cube 10^3, T=1000, delays 1-20, lr 1e-3, freeze_delays=False, mode {mode}")
    p = config0
    cfg = Algorithm(learning_rate=1e-3, freeze_delays=False)
    g, prof = run_gpu_epochs(gpu, p, cfg, profile=True, mode=mode)
    o32 = run_oracle_epochs(p, cfg, np.float32)
    o64 = run_oracle_epochs(p, cfg, np.float64)
    compare(g, o32, o64)
    note(f"  profile: {prof}")


@pytest.mark.parametrize("mode", [1, 2], ids=["per-step", "resident"])
def test_epoch_with_the_reference_ohara_drive(gpu, mode):
    """The reference's default drive (its assets/control_function_ohara.npy,
    kept as a fixture and tiled modulo its 2001 samples as control.rs:215-221
    does) over 2300 steps, so it wraps once: one epoch against the oracle."""
    note(f"case: blob of 300 voxels, O'Hara drive, T=2300, delays 1-15, mode {mode}")
    raw = np.load(os.path.join(os.path.dirname(__file__), "golden",
                               "control_function_ohara.npy"), allow_pickle=False)
    T = 2300
    p = synth.make_problem(synth.blob(77, 300), num_steps=T, max_delay=15, seed=77,
                           gain_sigma=0.1)
    p.control_function = synth.tiled_control_function(raw, T)
    p.measurements = oracle.simulate_f32(p.truth())
    cfg = Algorithm(learning_rate=1e-3, freeze_delays=False)
    g, _ = run_gpu_epochs(gpu, p, cfg, mode=mode)
    compare(g, run_oracle_epochs(p, cfg, np.float32), run_oracle_epochs(p, cfg, np.float64))


def blob_problem(seed, shape=(7, 6, 5), T=64, max_delay=12, beats=1, fill=0.6, gain_sigma=0.12):
    rng_u = synth.uniform(seed, 99, int(np.prod(shape))).reshape(shape)
    mask = rng_u < fill
    mask[0, 0, 0] = True
    g = synth.grid_geometry(mask, 2.5)
    offs = np.stack([np.array([0.0, 0.0, 0.0]) + 3.0 * b for b in range(beats)])
    return synth.make_problem(g, num_steps=T, max_delay=max_delay, seed=seed, num_beats=beats,
                              gain_sigma=gain_sigma, beat_offsets_mm=offs,
                              simulate=oracle.simulate_f32)


@pytest.mark.parametrize("case", [
    dict(name="textbook-maxreg", cfg=dict(learning_rate=0.5, freeze_delays=False,
                                          maximum_regularization_threshold=0.05)),
    dict(name="simple-diffreg", cfg=dict(learning_rate=0.5, freeze_delays=False,
                                         ap_derivative=APDerivative.Simple,
                                         difference_regularization_strength=1e-3)),
    dict(name="smoothness", cfg=dict(learning_rate=0.5, freeze_delays=False,
                                     smoothness_regularization_strength=0.1)),
    dict(name="frozen-gains", cfg=dict(learning_rate=0.5, freeze_delays=False, freeze_gains=True)),
    dict(name="frozen-both", cfg=dict(learning_rate=0.5, freeze_delays=True, freeze_gains=True)),
    dict(name="short-T", T=9, maxd=15, cfg=dict(learning_rate=0.5, freeze_delays=False)),
])
def test_epoch_small_cases(gpu, case):
    T = case.get("T", 64)
    maxd = case.get("maxd", 12)
    note(f"case: {case['name']} blob 7x6x5 seed 5, T={T}, max delay {maxd}")
    p = blob_problem(5, T=T, max_delay=maxd)
    cfg = Algorithm(**case["cfg"])
    g, _ = run_gpu_epochs(gpu, p, cfg, epochs=2, nbatch=2)
    o32 = run_oracle_epochs(p, cfg, np.float32, epochs=2, nbatch=2)
    o64 = run_oracle_epochs(p, cfg, np.float64, epochs=2, nbatch=2)
    # after the first update the parameters differ by rounding, so the second
    # epoch's states are compared with the tolerance
    exact = ("delays",)
    compare(g, o32, o64, exact=exact, close=CLOSE + ("system_states", "ap_outputs_now"))


@pytest.mark.parametrize("mode", [1, 2], ids=["per-step", "resident"])
def test_first_epoch_states_bit_exact_on_ragged_mask(gpu, mode):
    note(f"case: ragged blob 9x4x6 seed 8, T=80, one epoch, textbook, mode {mode}")
    p = blob_problem(8, shape=(9, 4, 6), T=80, max_delay=20, fill=0.45)
    cfg = Algorithm(learning_rate=0.2, freeze_delays=False)
    g, _ = run_gpu_epochs(gpu, p, cfg, mode=mode)
    o32 = run_oracle_epochs(p, cfg, np.float32)
    o64 = run_oracle_epochs(p, cfg, np.float64)
    compare(g, o32, o64)


def test_adam_first_update_within_1e4(gpu):
    """One Adam update (3 beats, batch_size 0, beat order [2, 0, 1]) against
    the oracle at the 1e-4 bar. The states are computed before the update, so
    they are bit-exact. Adam's first step moves every parameter by
    (lr / bs) * m1c / (sqrt(m2c) + eps), which is sign(d) to within 1e-7
    whenever |d| >> eps: an entry whose derivative the GPU and the oracle give
    with the same sign moves identically, so gains and coefficients must
    agree at 1e-4 there. Where |d| is within the derivative bar itself
    (|d_oracle| <= 1e-4 max|d_oracle|) the two FP32 summation orders may
    round it to opposite signs, and the parameter may then differ by at most
    2 lr / bs; the test shows how many entries that is."""
    note("case: 3 beats, batch_size 0, Adam, beat order [2,0,1], 1 epoch")
    p = blob_problem(3, T=48, beats=3)
    cfg = Algorithm(learning_rate=0.01, batch_size=0, optimizer=Optimizer.Adam, freeze_delays=False)
    order = [2, 0, 1]
    g, _ = run_gpu_epochs(gpu, p, cfg, beat_order=order, optimizer=1)
    o32 = run_oracle_epochs(p, cfg, np.float32, beat_order=order, optimizer=1)
    o64 = run_oracle_epochs(p, cfg, np.float64, beat_order=order, optimizer=1)
    assert g["derivatives_step"][0] == o32.d_step[0] == 2
    close = tuple(k for k in CLOSE if k not in ("gains", "coefs"))
    compare(g, o32, o64, exact=EXACT, close=close)
    step = cfg.learning_rate / (p.num_steps * p.num_beats)
    for name, dname in (("gains", "d_gains"), ("coefs", "d_coefs")):
        d_o = getattr(o32, dname).astype(np.float64)
        same = np.sign(g[dname]) == np.sign(d_o)
        ref = getattr(o32, name).astype(np.float64)
        diff = np.abs(g[name] - ref)
        scale = np.max(np.abs(ref))
        assert np.max(diff[same], initial=0.0) <= REL * scale, name
        flipped = ~same
        assert np.all(np.abs(d_o[flipped]) <= REL * np.max(np.abs(d_o))), name
        assert np.all(diff[flipped] <= 2 * step * (1 + 1e-5) + REL * scale), name
        note(f"  {name}: {int(flipped.sum())} of {flipped.size} entries with a derivative sign "
             f"at FP32 noise; worst same-sign error {np.max(diff[same]) / scale:.2e}")


def test_multibeat_batches_adam(gpu):
    """Two epochs of two Adam updates each (batch_size 2 over 3 beats). The
    bar is 1e-3, not 1e-4: every Adam step moves a parameter by about
    lr / bs = 1e-4 whatever the derivative's size, so a derivative at the FP32
    summation noise that the GPU and the oracle round to opposite signs leaves
    the parameter 2e-4 apart (7e-4 of the largest gain), and the later updates
    and states inherit it. test_adam_first_update_within_1e4 shows the same
    update at 1e-4 once those entries are accounted for."""
    note("case: 3 beats, batch_size 2, Adam, beat order [2,0,1], 2 epochs")
    p = blob_problem(3, T=48, beats=3)
    cfg = Algorithm(learning_rate=0.01, batch_size=2, optimizer=Optimizer.Adam, freeze_delays=False)
    order = [2, 0, 1]
    g, _ = run_gpu_epochs(gpu, p, cfg, epochs=2, beat_order=order, nbatch=4, optimizer=1)
    o32 = run_oracle_epochs(p, cfg, np.float32, epochs=2, beat_order=order, nbatch=4, optimizer=1)
    o64 = run_oracle_epochs(p, cfg, np.float64, epochs=2, beat_order=order, nbatch=4, optimizer=1)
    assert g["batch_index"] == o32.batch_index.value == 4
    assert g["derivatives_step"][0] == o32.d_step[0] == 5
    compare(g, o32, o64, exact=("delays",), close=CLOSE + ("system_states",), rel=1e-3)


def test_resident_and_per_step_paths_agree_bitwise(gpu):
    """The per-step launches (1), the resident kernel exchanging its state
    rows through agent-coherent stores and loads (2), and through agent-scope
    fences (3, the path of problems with 12 T N >= 2^32 bytes of states)."""
    from cardiotrust_amd.gpu import ResultsGPU
    note("case: shell 24^3 (several resident workgroups), simulate, modes 1 vs 2 vs 3")
    g0 = synth.ellipsoid_shell(shape=(24, 24, 24), outer=(12, 12, 12), inner=(6, 6, 6))
    p = synth.make_problem(g0, num_steps=150, max_delay=40, seed=9)
    outs = []
    for mode in (1, 2, 3):
        r = ResultsGPU(gpu, p)
        assert r.set_propagation(mode) == (mode != 1)
        m = r.simulate()
        x = r.update_from_gpu(["system_states", "ap_outputs_now", "ap_outputs_last"])
        r.close()
        outs.append((m, x))
    assert g0.num_voxels > 5000
    for other in outs[1:]:
        for k in ("system_states", "ap_outputs_now", "ap_outputs_last"):
            assert np.array_equal(outs[0][1][k].view(np.uint32), other[1][k].view(np.uint32)), k
        assert np.array_equal(outs[0][0], other[0])


def test_resident_long_delays_bitwise(gpu):
    """Slabs whose delays all fit a byte run their connections sorted by
    (delay, neighbour); a slab with any delay above 255 keeps connection order.
    Both kinds in one launch must equal the per-step path bit for bit."""
    from cardiotrust_amd.gpu import ResultsGPU
    note("case: shell 24^3, T=420, delays 256-400 on the first 1500 voxels, modes 1 vs 2 vs 3")
    g0 = synth.ellipsoid_shell(shape=(24, 24, 24), outer=(12, 12, 12), inner=(6, 6, 6))
    p = synth.make_problem(g0, num_steps=420, max_delay=40, seed=12)
    long = 256 + (synth.uniform(12, 40, 1500 * 26) * 145).astype(np.int32).reshape(1500, 26)
    p.delays[:1500] = np.where(g0.neighbours[:1500] >= 0, long, p.delays[:1500])
    assert p.delays.max() > 255 and p.delays[1500:].max() <= 40
    outs = []
    for mode in (1, 2, 3):
        r = ResultsGPU(gpu, p)
        assert r.set_propagation(mode) == (mode != 1)
        m = r.simulate()
        x = r.update_from_gpu(["system_states", "ap_outputs_now", "ap_outputs_last"])
        r.close()
        outs.append((m, x))
    for other in outs[1:]:
        for k in ("system_states", "ap_outputs_now", "ap_outputs_last"):
            assert np.array_equal(outs[0][1][k].view(np.uint32), other[1][k].view(np.uint32)), k
        assert np.array_equal(outs[0][0], other[0])


def test_simulate_matches_oracle(gpu):
    from cardiotrust_amd.gpu import ResultsGPU
    note("case: simulate blob seed 4, 2 beats")
    p = blob_problem(4, T=70, beats=2)
    r = ResultsGPU(gpu, p)
    m = r.simulate()
    x = r.update_from_gpu(["system_states"])["system_states"]
    r.close()
    o = oracle.OracleResults(p, np.float32)
    om = o.simulate()
    assert np.array_equal(x.view(np.uint32), o.system_states.view(np.uint32))
    assert rel_inf(m, om) <= REL


def test_empty_model(gpu):
    note("case: empty model (no voxels)")
    g0 = synth.grid_geometry(np.zeros((2, 2, 2), dtype=bool), 2.5)
    p = synth.make_problem(g0, num_steps=16, max_delay=3, seed=0)
    p.measurements = synth.uniform(0, 7, 16 * 108).reshape(1, 16, 108).astype(np.float32)
    cfg = Algorithm(learning_rate=1.0, freeze_delays=False)
    g, _ = run_gpu_epochs(gpu, p, cfg)
    o32 = run_oracle_epochs(p, cfg, np.float32)
    for k in ("loss_mse", "loss_batch", "residuals"):
        assert rel_inf(g[k], getattr(o32, k)) <= REL, k


def test_invalid_inputs_are_rejected(gpu):
    from cardiotrust_amd.gpu import Buffer, CTError, EpochKernel, ResultsGPU
    note("case: invalid inputs")
    p = blob_problem(6, T=20)
    bad = blob_problem(6, T=20)
    conn = np.argwhere(bad.geometry.neighbours >= 0)[0]
    bad.delays[conn[0], conn[1]] = 0
    with pytest.raises(CTError, match="delay < 1"):
        ResultsGPU(gpu, bad)
    bad = blob_problem(6, T=20)
    bad.output_state_indices[1, 5] = 7 if bad.output_state_indices[1, 5] != 7 else 8
    with pytest.raises(CTError, match="26-neighbour"):
        ResultsGPU(gpu, bad)
    r = ResultsGPU(gpu, p)
    buf = Buffer(gpu, p.measurements)
    with pytest.raises(CTError, match="do not match"):
        EpochKernel(gpu, r, buf, Algorithm(), 3 * p.num_voxels + 3, p.num_sensors, p.num_steps)
    k = EpochKernel(gpu, r, buf, Algorithm(freeze_delays=False), 3 * p.num_voxels, p.num_sensors,
                    p.num_steps)
    with pytest.raises(CTError, match="permutation"):
        k.set_beat_order([1])
    k.execute()
    with pytest.raises(CTError, match="batch index"):
        k.execute()  # only one batch-metric entry
    k.close()
    buf.close()
    r.close()


def test_simulate_with_measurement_noise(gpu):
    """ct_results_simulate_noisy = the noise-free simulation plus the seeded
    noise that synth.measurement_noise mirrors on the host (FP64 Box-Muller
    on both sides: equal up to one FP32 ulp where the two libm differ)."""
    from cardiotrust_amd.gpu import CTError, ResultsGPU
    note("case: simulate blob seed 4, 2 beats, with per-sensor noise, seed 99")
    p = blob_problem(4, T=70, beats=2)
    r = ResultsGPU(gpu, p)
    clean = r.simulate()
    sd = (0.01 * (1.0 + np.arange(p.num_sensors) % 5)).astype(np.float32)
    noisy = r.simulate(noise_std=sd, seed=99)
    want = (clean + synth.measurement_noise(99, sd, p.num_beats, p.num_steps)).astype(np.float32)
    ulp = np.abs(noisy.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    assert ulp.max() <= 1 and np.mean(ulp == 0) > 0.999
    again = r.simulate(noise_std=sd, seed=99)
    assert np.array_equal(again, noisy)
    assert np.array_equal(r.simulate(noise_std=np.zeros_like(sd), seed=5), clean)
    for bad in (-1.0, np.nan, np.inf):
        sd2 = sd.copy()
        sd2[3] = bad
        with pytest.raises(CTError):
            r.simulate(noise_std=sd2, seed=1)
    r.close()


def test_measurement_matrix_on_device_equals_the_oracle(gpu):
    """ct_measurement_matrix runs measurement.rs:66-123 in the reference's f32
    operations (correctly rounded division and square root on gfx950), so it
    equals the oracle's restatement bit for bit: a blob with two beats and the
    120k-voxel shell with one."""
    import time
    from cardiotrust_amd.gpu import CTError, measurement_matrix
    for g, offs in ((synth.blob(9, 500), np.array([[0.0, 0.0, 0.0], [12.5, -3.0, 4.0]])),
                    (synth.ellipsoid_shell(), np.zeros((1, 3)))):
        note(f"case: Biot-Savart matrix, {g.num_voxels} voxels, {len(offs)} beats")
        s = synth.sensor_grid(g, num_beats=len(offs), beat_offsets_mm=offs)
        t0 = time.perf_counter()
        H = measurement_matrix(gpu, g.positions_mm, s.positions_mm, s.orientations, offs,
                               g.voxel_size_mm)
        dt = time.perf_counter() - t0
        want = oracle.measurement_matrix_f32(g.positions_mm, s.positions_mm, s.orientations, offs,
                                             g.voxel_size_mm)
        note(f"  {H.nbytes / 1e6:.0f} MB in {dt * 1e3:.1f} ms (with the copy back)")
        assert np.array_equal(H.view(np.uint32), want.view(np.uint32))
    with pytest.raises(CTError):
        measurement_matrix(gpu, np.zeros((2, 3)), np.ones((1, 3)), np.eye(3)[:1], np.zeros((1, 3)),
                           -1.0)


def test_model_connections_on_device_equal_the_oracle(gpu):
    """ct_model_connections builds output_state_indices and the delays in
    samples with the reference's f32 operations: equal to the oracle bit for
    bit on a ragged blob (with per-cell velocities) and the 120k-voxel shell;
    a sub-sample delay and malformed numbers are errors."""
    import time
    from cardiotrust_amd.gpu import CTError, model_connections
    for g in (synth.blob(6, 900), synth.ellipsoid_shell()):
        note(f"case: model connections, grid {g.shape}, {g.num_voxels} voxels")
        numbers = np.where(g.numbers >= 0, 3 * g.numbers, -1).astype(np.int32)
        idx = np.indices(g.shape).transpose(1, 2, 3, 0).astype(np.float64)
        pos = ((idx + 0.5) * g.voxel_size_mm).astype(np.float32)
        vel = (0.6 + 0.4 * synth.uniform(6, 3, int(np.prod(g.shape)))).astype(np.float32)
        t0 = time.perf_counter()
        osi, dly = model_connections(gpu, numbers, pos, vel.reshape(g.shape), 2000.0, g.num_voxels)
        note(f"  {(time.perf_counter() - t0) * 1e3:.1f} ms with the copies")
        o_osi, o_dly = oracle.model_connections_f32(numbers, pos, vel.reshape(g.shape), 2000.0,
                                                    g.num_voxels)
        assert np.array_equal(osi, o_osi)
        assert np.array_equal(dly.view(np.uint32), o_dly.view(np.uint32))
    g = synth.cube(3, voxel_size_mm=0.1)
    numbers = (3 * g.numbers).astype(np.int32)
    pos = ((np.indices(g.shape).transpose(1, 2, 3, 0) + 0.5) * 0.1).astype(np.float32)
    with pytest.raises(CTError):
        model_connections(gpu, numbers, pos, np.ones(g.shape, np.float32), 2000.0, g.num_voxels)
    bad = numbers.copy()
    bad[0, 0, 0] = 1  # not a multiple of 3
    with pytest.raises(CTError):
        model_connections(gpu, bad, pos, np.ones(g.shape, np.float32), 2.0, g.num_voxels)


def test_spherical_states_on_device(gpu):
    """ct_results_spherical over the device states of a simulation against the
    oracle's restatement on the same states: magnitudes, the argmax-based
    activation times and maxima bit for bit; theta and phi within 4 FP32 ulp
    (acosf / atan2f of the device library against the C library's)."""
    from cardiotrust_amd.gpu import ResultsGPU
    note("case: spherical states of a 2-beat blob simulation and of the 120k-voxel shell")
    for p in (blob_problem(4, T=70, beats=2),
              synth.make_problem(synth.ellipsoid_shell(), num_steps=300, max_delay=40, seed=3)):
        r = ResultsGPU(gpu, p)
        r.simulate()
        x = r.update_from_gpu(["system_states"])["system_states"]
        got = r.spherical(2000.0, full=True)
        r.close()
        want = oracle.spherical_f32(x, 2000.0)
        for k in ("magnitude", "magnitude_max", "activation_time_ms"):
            assert np.array_equal(got[k].view(np.uint32), want[k].view(np.uint32)), k
        for k in ("theta", "phi", "theta_max", "phi_max"):
            a, b = got[k], want[k]
            assert np.array_equal(np.isnan(a), np.isnan(b)), k
            ok = ~np.isnan(a)
            ulp = np.abs(a[ok].view(np.int32).astype(np.int64) - b[ok].view(np.int32).astype(np.int64))
            assert ulp.max(initial=0) <= 4, (k, ulp.max())
        note(f"  N={p.num_voxels}: activation times {np.unique(got['activation_time_ms']).size} distinct")


def test_wait_timeout_aborts_and_falls_back_to_per_step(gpu):
    """A failure forced from the host through the kernel's own bounded wait: a
    wait limit of 0 s makes a resident launch time out at the first wait that
    is not already satisfied. Every workgroup then leaves at its next wait, the
    call reports CT_ERR_DEVICE, and the results object runs the per-step
    kernels, whose states equal the per-step path's bit for bit. (A launch in
    which no wait has to wait completes normally; one of three attempts must
    time out.)"""
    from cardiotrust_amd.gpu import CTError, ResultsGPU
    note("case: 24^3 shell, T=600, resident wait limit 0 s, then the per-step fallback")
    g0 = synth.ellipsoid_shell(shape=(24, 24, 24), outer=(12, 12, 12), inner=(6, 6, 6))
    p = synth.make_problem(g0, num_steps=600, max_delay=40, seed=14)
    ref = ResultsGPU(gpu, p)
    ref.set_propagation(1)
    m_ref = ref.simulate()
    x_ref = ref.update_from_gpu(["system_states"])["system_states"]
    ref.close()
    r = ResultsGPU(gpu, p)
    assert r.set_propagation(0)
    r.set_wait_limit(0.0)
    timed_out = False
    for _ in range(3):
        try:
            r.simulate()
        except CTError as e:
            assert "timed out" in str(e)
            timed_out = True
            break
    assert timed_out
    assert not r.set_propagation(0)  # the resident path is off for this problem now
    r.set_wait_limit(2.0)
    m = r.simulate()
    x = r.update_from_gpu(["system_states"])["system_states"]
    r.close()
    assert np.array_equal(x.view(np.uint32), x_ref.view(np.uint32))
    assert np.array_equal(m, m_ref)
    with pytest.raises(CTError):
        ResultsGPU(gpu, p).set_wait_limit(-1.0)


def test_wait_timeout_in_epoch_leaves_parameters(gpu):
    """The same host-forced timeout inside an epoch: the derivatives of a
    timed-out propagation are not applied. The SGD update and the delay roll
    read the results' abort word and skip, so after the CT_ERR_DEVICE the
    gains, coefficients, delays and the batch index are those from before the
    call. With the limit restored, the next epoch runs on the per-step path and
    equals a per-step epoch from the same parameters bit for bit."""
    from cardiotrust_amd.gpu import Buffer, CTError, EpochKernel, ResultsGPU
    note("case: 24^3 shell, T=300, epoch with resident wait limit 0 s, then the per-step epoch")
    g0 = synth.ellipsoid_shell(shape=(24, 24, 24), outer=(12, 12, 12), inner=(6, 6, 6))
    p = synth.make_problem(g0, num_steps=300, max_delay=40, seed=15)
    truth = ResultsGPU(gpu, p.truth())
    truth.set_propagation(1)
    p.measurements = truth.simulate()
    truth.close()
    cfg = Algorithm(learning_rate=0.5, freeze_delays=False)
    params = ("gains", "coefs", "delays")

    def epoch_object(mode):
        r = ResultsGPU(gpu, p, num_batch_entries=4)
        r.set_propagation(mode)
        buf = Buffer(gpu, p.measurements)
        return r, buf, EpochKernel(gpu, r, buf, cfg, 3 * p.num_voxels, p.num_sensors, p.num_steps)

    r, buf, k = epoch_object(0)
    assert r.set_propagation(0)
    r.set_wait_limit(0.0)
    before = r.update_from_gpu(list(params))
    index = k.batch_index()
    ok_epochs, failed = 0, False
    for _ in range(3):
        try:
            k.execute()
        except CTError as e:
            assert "timed out" in str(e), e
            failed = True
            break
        ok_epochs += 1
        before = r.update_from_gpu(list(params))
        index = k.batch_index()
    assert failed
    after = r.update_from_gpu(list(params))
    for name in params:
        assert np.array_equal(before[name], after[name]), name
    assert k.batch_index() == index
    assert not r.set_propagation(0)
    r.set_wait_limit(2.0)
    k.execute()
    got = r.update_from_gpu(list(params) + ["system_states", "loss_batch"])
    k.close(), buf.close(), r.close()
    r, buf, k = epoch_object(1)
    for _ in range(ok_epochs + 1):
        k.execute()
    want = r.update_from_gpu(list(params) + ["system_states", "loss_batch"])
    k.close(), buf.close(), r.close()
    for name in want:
        assert np.array_equal(got[name], want[name]), name


@pytest.mark.parametrize("mode", [1, 2], ids=["per-step", "resident"])
def test_more_than_128_sensors(gpu, mode):
    """243 sensors (a 9 x 9 grid of 3-axis magnetometers) pad to two 128-row
    sensor tiles: H·Xᵀ runs two sensor tiles (only the first writes xT), Hᵀ·R
    eight k tiles. Two epochs against the oracle; states and delays bit-exact
    in the first."""
    note(f"case: blob seed 11, 243 sensors, T=90, 2 epochs, mode {mode}")
    rng_u = synth.uniform(11, 99, 7 * 6 * 5).reshape(7, 6, 5)
    mask = rng_u < 0.6
    mask[0, 0, 0] = True
    g = synth.grid_geometry(mask, 2.5)
    p = synth.make_problem(g, num_steps=90, max_delay=14, seed=11, gain_sigma=0.12,
                           simulate=oracle.simulate_f32, sensors_per_axis=9)
    assert p.num_sensors == 243
    cfg = Algorithm(learning_rate=0.3, freeze_delays=False, maximum_regularization_threshold=0.05)
    g1, _ = run_gpu_epochs(gpu, p, cfg, mode=mode)
    o32 = run_oracle_epochs(p, cfg, np.float32)
    o64 = run_oracle_epochs(p, cfg, np.float64)
    compare(g1, o32, o64)
    g2, _ = run_gpu_epochs(gpu, p, cfg, epochs=2, nbatch=2, mode=mode)
    o32 = run_oracle_epochs(p, cfg, np.float32, epochs=2, nbatch=2)
    o64 = run_oracle_epochs(p, cfg, np.float64, epochs=2, nbatch=2)
    compare(g2, o32, o64, exact=("delays",), close=CLOSE + ("system_states",))
