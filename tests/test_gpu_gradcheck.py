"""Reverse-mode gradient against FP64 central finite differences at full size
(32^3 cube, N = 32,768 voxels, T = 2000 steps, 108 sensors).

The FP64 objective (ct_results_objective_f64) is first pinned to the CPU
oracle's FP64 objective, unperturbed and with single gains / coefficients
shifted in FP64. Then, for checkpoint intervals K = 16, 64, 256:
  * the GPU gradient is the same bit for bit for every K;
  * (J(p + h) - J(p - h)) / 2h with h = 1e-6 in FP64 agrees with the GPU's
    FP32 reverse-mode gradient on seeded draws of 64 gains and 64
    coefficients: normwise (1e-3 of max |g|) on a uniform draw, and per entry
    (1e-3 relative) on a draw among the entries above the FD noise floor.
The truncated mode (the reference's own derivation) is compared with the
oracle's run_epoch derivatives on the same 32^3 parameters over 200 steps.
"""
import sys
import time

import numpy as np
import pytest

from cardiotrust_amd import Algorithm, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

H_FD = 1e-6
PICKS = 64


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel_inf(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def gpu():
    from cardiotrust_amd.gpu import GPU
    g = GPU(0)
    yield g
    g.close()


def blob(seed, shape=(6, 5, 5), T=60, maxd=10):
    mask = synth.uniform(seed, 99, int(np.prod(shape))).reshape(shape) < 0.7
    mask[0, 0, 0] = True
    g = synth.grid_geometry(mask, 2.5)
    return synth.make_problem(g, num_steps=T, max_delay=maxd, seed=seed, gain_sigma=0.25,
                              simulate=oracle.simulate_f32)


def test_objective_f64_matches_oracle(gpu):
    from cardiotrust_amd.gpu import Buffer, ResultsGPU
    note("case: FP64 objective vs oracle FP64, blob seeds 11/12, unperturbed and shifted")
    for seed, theta in ((11, 0.05), (12, 1.01)):
        p = blob(seed)
        cfg = Algorithm(maximum_regularization_threshold=theta)
        r = ResultsGPU(gpu, p)
        buf = Buffer(gpu, p.measurements)
        o = oracle.OracleResults(p, np.float64)
        J0 = r.objective_f64(buf, 0, cfg)
        assert J0 == pytest.approx(o.adjoint(cfg)[0], rel=1e-10)
        rng = np.random.default_rng(seed)
        conn = np.argwhere(p.output_state_indices >= 0)
        for i, g in conn[rng.choice(len(conn), 3, replace=False)]:
            d = 1e-3
            J = r.objective_f64(buf, 0, cfg, param=1, index=int(i) * 78 + int(g), delta=d)
            o.gains[i, g] += d
            assert J == pytest.approx(o.adjoint(cfg)[0], rel=1e-10), (i, g)
            o.gains[i, g] -= d
        vox = np.argwhere(p.geometry.neighbours >= 0)
        for v, dd in vox[rng.choice(len(vox), 3, replace=False)]:
            d = -2e-3
            J = r.objective_f64(buf, 0, cfg, param=2, index=int(v) * 26 + int(dd), delta=d)
            o.coefs[v, dd] += d
            assert J == pytest.approx(o.adjoint(cfg)[0], rel=1e-10), (v, dd)
            o.coefs[v, dd] -= d
        buf.close()
        r.close()


@pytest.fixture(scope="module")
def cube32(gpu):
    from cardiotrust_amd.gpu import ResultsGPU
    p = synth.config2(num_steps=2000)
    truth = ResultsGPU(gpu, p.truth())
    p.measurements = truth.simulate()
    truth.close()
    return p


def central_differences(r, buf, cfg, param, indices):
    out = []
    for idx in indices:
        jp = r.objective_f64(buf, 0, cfg, param, int(idx), H_FD)
        jm = r.objective_f64(buf, 0, cfg, param, int(idx), -H_FD)
        out.append((jp - jm) / (2 * H_FD))
    return np.array(out)


def test_adjoint_matches_fp64_central_differences_at_32k(gpu, cube32):
    """Two seeded draws of 64 gains and 64 coefficients (existing connections):
    (a) uniformly over all of them: most lie where the pulse from the driven
        corner has decayed and the true gradient is far below the FD noise
        floor (~ J eps / h ~ 1e-9), so the bar is normwise:
        max |g - fd| <= 1e-3 max |g| over the whole gradient;
    (b) uniformly over the resolvable ones, |g| >= 1e4 x that floor
        (J eps / h), where FD noise is below 1e-4 relative: per entry
        |g - fd| <= 1e-3 |fd|."""
    from cardiotrust_amd.gpu import Buffer, ResultsGPU
    p = cube32
    note(f"case: 32^3 cube N={p.num_voxels}, T={p.num_steps}, K in 16/64/256, "
         f"2 x ({PICKS} gains + {PICKS} coefs) FP64 central differences, h={H_FD}")
    cfg = Algorithm(maximum_regularization_threshold=0.05)
    r = ResultsGPU(gpu, p)
    buf = Buffer(gpu, p.measurements)
    grads = {}
    for K in (16, 64, 256):
        t0 = time.perf_counter()
        J = r.forward(buf, 0, cfg, checkpoint_interval=K)
        out = r.backward(buf, 0, cfg, checkpoint_interval=K)
        grads[K] = (out["d_gains"], out["d_coefs"])
        note(f"  K={K}: J={J:.9e}, forward+backward {time.perf_counter() - t0:.2f} s")
    for K in (64, 256):
        assert np.array_equal(grads[K][0].view(np.uint32), grads[16][0].view(np.uint32)), K
        assert np.array_equal(grads[K][1].view(np.uint32), grads[16][1].view(np.uint32)), K
    dG, dA = grads[16]
    J64 = r.objective_f64(buf, 0, cfg)
    assert J == pytest.approx(J64, rel=1e-5)
    rng = np.random.default_rng(2)
    gains_all = np.flatnonzero(p.output_state_indices.ravel() >= 0)
    coefs_all = np.flatnonzero(p.geometry.neighbours.ravel() >= 0)
    t0 = time.perf_counter()
    for name, param, grad, pool in (("gains", 1, dG.ravel(), gains_all),
                                    ("coefs", 2, dA.ravel(), coefs_all)):
        scale = float(np.max(np.abs(grad)))
        uni = rng.choice(pool, PICKS, replace=False)
        fd = central_differences(r, buf, cfg, param, uni)
        err_a = float(np.max(np.abs(grad[uni] - fd)))
        floor = 1e4 * abs(J64) * np.finfo(np.float64).eps / H_FD
        resolvable = pool[np.abs(grad[pool]) >= floor]
        res = rng.choice(resolvable, min(PICKS, resolvable.size), replace=False)
        fd_r = central_differences(r, buf, cfg, param, res)
        err_b = float(np.max(np.abs(grad[res] - fd_r) / np.abs(fd_r)))
        note(f"  {name}: max|g| {scale:.3e}; uniform draw max|g-fd| {err_a:.2e} "
             f"(= {err_a / scale:.2e} max|g|); {res.size} of {resolvable.size} above "
             f"{floor:.1e}: "
             f"max rel err {err_b:.2e} ({time.perf_counter() - t0:.1f} s)")
        assert res.size == PICKS
        assert err_a <= 1e-3 * scale, name
        assert err_b <= 1e-3, name
    buf.close()
    r.close()


def test_truncated_mode_is_the_reference_derivation_at_32k(gpu, cube32):
    from cardiotrust_amd.gpu import Buffer, ResultsGPU
    T = 200
    note(f"case: 32^3 cube, backward(truncate=1) vs oracle run_epoch derivatives, T={T}")
    full = cube32
    p = synth.Problem(full.geometry, full.num_sensors, T, 1, full.gains,
                      full.output_state_indices, full.coefs, full.delays, full.initial_delays,
                      full.control_matrix, full.control_function[:T], full.measurement_matrix,
                      np.ascontiguousarray(full.measurements[:, :T]))
    cfg = Algorithm(learning_rate=0.0, freeze_delays=False, maximum_regularization_threshold=0.05)
    r = ResultsGPU(gpu, p)
    buf = Buffer(gpu, p.measurements)
    r.forward(buf, 0, cfg, checkpoint_interval=16)
    out = r.backward(buf, 0, cfg, truncate=True, checkpoint_interval=16)
    buf.close()
    r.close()
    o = oracle.OracleResults(p, np.float32)
    o.run_epoch(cfg)
    assert rel_inf(out["d_gains"], o.d_gains) <= 1e-4
    assert rel_inf(out["d_coefs"], o.d_coefs) <= 1e-4
