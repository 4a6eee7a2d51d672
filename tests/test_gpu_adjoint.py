"""Reverse-mode gradient on the GPU against the oracle's FP64 adjoint (itself
pinned by central finite differences in test_oracle_adjoint.py).

Bars: J within 1e-5 relative; dJ/dgains and dJ/dcoefs within 1e-3 relative
(||gpu - f64||_inf / ||f64||_inf, FP32 accumulation over T steps); results
independent of the checkpoint interval bit for bit; the truncated mode equal
to the reference derivation within 1e-4.
"""
import sys

import numpy as np
import pytest

from cardiotrust_amd import Algorithm, synth
from oracle import oracle

pytestmark = pytest.mark.gpu


def rel_inf(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    from cardiotrust_amd.gpu import GPU
    g = GPU(0)
    yield g
    g.close()


def blob(seed, shape=(6, 5, 5), T=60, maxd=10, fill=0.7, sigma=0.25):
    mask = synth.uniform(seed, 99, int(np.prod(shape))).reshape(shape) < fill
    mask[0, 0, 0] = True
    g = synth.grid_geometry(mask, 2.5)
    return synth.make_problem(g, num_steps=T, max_delay=maxd, seed=seed, gain_sigma=sigma,
                              simulate=oracle.simulate_f32)


def gpu_grad(gpu, p, cfg, K, truncate=False, mode=0):
    from cardiotrust_amd.gpu import Buffer, ResultsGPU
    r = ResultsGPU(gpu, p)
    r.set_propagation(mode)
    buf = Buffer(gpu, p.measurements)
    J = r.forward(buf, 0, cfg, checkpoint_interval=K)
    out = r.backward(buf, 0, cfg, truncate=truncate, checkpoint_interval=K)
    buf.close()
    r.close()
    return J, out["d_gains"], out["d_coefs"]


@pytest.mark.parametrize("seed,theta", [(1, 0.05), (2, 1.01), (3, 0.2)])
def test_adjoint_matches_oracle(gpu, seed, theta):
    note(f"case: adjoint blob seed {seed}, theta {theta}")
    p = blob(seed)
    cfg = Algorithm(maximum_regularization_threshold=theta)
    J64, dG64, dA64 = oracle.OracleResults(p, np.float64).adjoint(cfg)
    J, dG, dA = gpu_grad(gpu, p, cfg, K=16)
    note(f"  J {J:.6e} vs {J64:.6e}; dG err {rel_inf(dG, dG64):.2e}; dA err {rel_inf(dA, dA64):.2e}")
    assert J == pytest.approx(J64, rel=1e-5)
    assert rel_inf(dG, dG64) <= 1e-3
    assert rel_inf(dA, dA64) <= 1e-3


def test_checkpoint_interval_does_not_change_the_gradient(gpu):
    note("case: adjoint checkpoint intervals 1, 7, 16, 64, 256 (T=100)")
    p = blob(4, T=100)
    cfg = Algorithm(maximum_regularization_threshold=0.1)
    ref = None
    for K in (1, 7, 16, 64, 256):
        J, dG, dA = gpu_grad(gpu, p, cfg, K=K)
        if ref is None:
            ref = (J, dG, dA)
        else:
            assert J == ref[0]
            assert np.array_equal(dG.view(np.uint32), ref[1].view(np.uint32)), K
            assert np.array_equal(dA.view(np.uint32), ref[2].view(np.uint32)), K


def test_truncated_backward_is_the_reference_derivation(gpu):
    note("case: backward(truncate=1) vs oracle run_epoch derivatives")
    p = blob(5)
    cfg = Algorithm(learning_rate=0.0, freeze_delays=False, maximum_regularization_threshold=0.1)
    _, dG, dA = gpu_grad(gpu, p, cfg, K=16, truncate=True)
    o = oracle.OracleResults(p, np.float32)
    o.run_epoch(cfg)
    assert rel_inf(dG, o.d_gains) <= 1e-4
    assert rel_inf(dA, o.d_coefs) <= 1e-4


def test_adjoint_epoch_update(gpu):
    from cardiotrust_amd.gpu import Buffer, EpochKernel, ResultsGPU
    note("case: epoch with the reverse-mode gradient, SGD update")
    p = blob(6)
    lr = 0.3
    cfg = Algorithm(learning_rate=lr, freeze_delays=False, maximum_regularization_threshold=0.1)
    r = ResultsGPU(gpu, p)
    buf = Buffer(gpu, p.measurements)
    k = EpochKernel(gpu, r, buf, cfg, 3 * p.num_voxels, p.num_sensors, p.num_steps)
    k.set_gradient(1, 16)
    k.execute()
    got = r.update_from_gpu(["gains", "coefs", "delays", "d_gains", "d_coefs"])
    k.close()
    buf.close()
    r.close()
    J64, dG64, dA64 = oracle.OracleResults(p, np.float64).adjoint(cfg)
    assert rel_inf(got["d_gains"], dG64) <= 1e-3
    assert rel_inf(got["d_coefs"], dA64) <= 1e-3
    # same SGD step and delay roll as the reference update (update.rs:27-228)
    o = oracle.OracleResults(p, np.float32)
    o.d_gains[:] = got["d_gains"]
    o.d_coefs[:] = got["d_coefs"]
    scale = np.float32(lr) / np.float32(p.num_steps)
    want_g = (o.gains - scale * o.d_gains).astype(np.float32)
    assert np.array_equal(got["gains"], want_g)
    want_a = (o.coefs - scale * o.d_coefs).astype(np.float32)
    roll = (want_a > np.float32(1 - 1e-4)) | (want_a < np.float32(1e-4))
    assert np.array_equal(got["coefs"][~roll], want_a[~roll])


def test_adjoint_resident_and_per_step_forward_agree(gpu):
    note("case: adjoint after resident vs per-step forward")
    p = blob(7, T=50)
    cfg = Algorithm(maximum_regularization_threshold=0.1)
    a = gpu_grad(gpu, p, cfg, K=16, mode=1)
    b = gpu_grad(gpu, p, cfg, K=16, mode=2)
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_resident_sweep_matches_per_step_sweep_on_many_workgroups(gpu):
    note("case: adjoint on a 24^3 shell (several resident workgroups), mode 1 vs 2 vs 3")
    g0 = synth.ellipsoid_shell(shape=(24, 24, 24), outer=(12, 12, 12), inner=(6, 6, 6))
    p = synth.make_problem(g0, num_steps=120, max_delay=40, seed=9)
    from cardiotrust_amd.gpu import ResultsGPU
    truth = ResultsGPU(gpu, p.truth())
    truth.set_propagation(1)
    p.measurements = truth.simulate()
    truth.close()
    cfg = Algorithm(maximum_regularization_threshold=0.05)
    a = gpu_grad(gpu, p, cfg, K=16, mode=1)
    assert g0.num_voxels > 5000
    for mode in (2, 3):  # coherent row exchange, and agent-scope fences
        b = gpu_grad(gpu, p, cfg, K=16, mode=mode)
        assert a[0] == b[0]
        assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
        assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))



def test_resident_sweep_long_delays_bitwise(gpu):
    """Sweep slabs whose delays all fit a byte run their items sorted by
    (delay, consumer); a slab with a delay above 255 keeps item order. Both in
    one launch must equal the per-step sweep bit for bit."""
    note("case: adjoint on a 24^3 shell, T=420, delays 256-400 on the first 1500 voxels, "
         "mode 1 vs mode 2")
    g0 = synth.ellipsoid_shell(shape=(24, 24, 24), outer=(12, 12, 12), inner=(6, 6, 6))
    p = synth.make_problem(g0, num_steps=420, max_delay=40, seed=13)
    long = 256 + (synth.uniform(13, 40, 1500 * 26) * 145).astype(np.int32).reshape(1500, 26)
    p.delays[:1500] = np.where(g0.neighbours[:1500] >= 0, long, p.delays[:1500])
    from cardiotrust_amd.gpu import ResultsGPU
    truth = ResultsGPU(gpu, p.truth())
    truth.set_propagation(1)
    p.measurements = truth.simulate()
    truth.close()
    p.gains = (p.gains * 1.1).astype(np.float32)  # a non-zero residual
    cfg = Algorithm(maximum_regularization_threshold=0.05)
    a = gpu_grad(gpu, p, cfg, K=64, mode=1)
    b = gpu_grad(gpu, p, cfg, K=64, mode=2)
    assert a[0] == b[0] and a[0] > 0
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32))
    assert np.abs(a[1]).max() > 0


@pytest.mark.parametrize("sensors_per_axis", [6, 9])
def test_reverse_mode_forcing_equals_the_mapped_residuals(gpu, sensors_per_axis):
    """The reverse-mode epoch writes Hᵀ·r by step (E[t][i]) where at most 112
    sensors allow it, else by state (eT[i][t]) and transposes; the reference
    derivation always writes eT. From the same start both epochs have the
    same states and residuals, so the downloaded mapped residuals (step T-1)
    must agree bit for bit. T = 150 leaves a partial 64-step tile; N is not a
    multiple of the 128-state tile."""
    from cardiotrust_amd.gpu import Buffer, EpochKernel, ResultsGPU
    p = blob(21, shape=(7, 6, 5), T=150, maxd=12)
    if sensors_per_axis != 6:
        g = synth.grid_geometry(synth.uniform(21, 99, 210).reshape(7, 6, 5) < 0.7, 2.5)
        p = synth.make_problem(g, num_steps=150, max_delay=12, seed=21, gain_sigma=0.25,
                               simulate=oracle.simulate_f32, sensors_per_axis=sensors_per_axis)
    note(f"case: forcing layout, {p.num_sensors} sensors, N={p.num_voxels}, T={p.num_steps}")
    assert (3 * p.num_voxels) % 128 != 0
    cfg = Algorithm(learning_rate=0.1, maximum_regularization_threshold=0.05)
    got = {}
    for mode in ("reference", "adjoint"):
        r = ResultsGPU(gpu, p)
        buf = Buffer(gpu, p.measurements)
        k = EpochKernel(gpu, r, buf, cfg, 3 * p.num_voxels, p.num_sensors, p.num_steps)
        if mode == "adjoint":
            k.set_gradient(1, 16)
        k.execute()
        got[mode] = r.update_from_gpu(["mapped_residuals", "residuals", "system_states"])
        k.close()
        buf.close()
        r.close()
    ref, adj = got["reference"], got["adjoint"]
    assert np.array_equal(ref["system_states"].view(np.uint32), adj["system_states"].view(np.uint32))
    assert np.array_equal(ref["residuals"].view(np.uint32), adj["residuals"].view(np.uint32))
    assert np.any(ref["mapped_residuals"] != 0)
    assert np.array_equal(ref["mapped_residuals"].view(np.uint32),
                          adj["mapped_residuals"].view(np.uint32))
