"""The oracle's reverse-mode gradient against float64 central finite
differences of its own objective (what fixes an adjoint: any dropped term,
wrong sign, delay index or transposed gain shows up as a mismatch)."""
import numpy as np
import pytest

from cardiotrust_amd import Algorithm, synth
from oracle import oracle


def problem(shape, T, maxd, seed, sigma=0.3):
    g = synth.grid_geometry(np.ones(shape, dtype=bool), 2.5)
    p = synth.make_problem(g, num_steps=T, max_delay=maxd, seed=seed, gain_sigma=sigma,
                           simulate=oracle.simulate_f32)
    return p


@pytest.mark.parametrize("shape,T,maxd,seed,theta", [
    ((3, 3, 3), 40, 6, 1, 0.05),   # maximum regularisation active
    ((2, 3, 4), 30, 9, 2, 1.01),   # mse only
    ((4, 1, 2), 50, 3, 3, 0.2),
])
def test_adjoint_matches_central_differences(shape, T, maxd, seed, theta):
    p = problem(shape, T, maxd, seed)
    cfg = Algorithm(maximum_regularization_threshold=theta, mse_strength=1.0)
    r = oracle.OracleResults(p, np.float64)
    J, dG, dA = r.adjoint(cfg)
    assert J > 0
    rng = np.random.default_rng(seed)
    h = 1e-6
    conn = np.argwhere(p.output_state_indices >= 0)
    picks = conn[rng.choice(len(conn), 16, replace=False)]
    fd_g, ad_g = [], []
    for i, g in picks:
        base = r.gains[i, g]
        r.gains[i, g] = base + h
        jp = r.adjoint(cfg)[0]
        r.gains[i, g] = base - h
        jm = r.adjoint(cfg)[0]
        r.gains[i, g] = base
        fd_g.append((jp - jm) / (2 * h))
        ad_g.append(dG[i, g])
    vox = np.argwhere(p.geometry.neighbours >= 0)
    vpicks = vox[rng.choice(len(vox), 16, replace=False)]
    fd_a, ad_a = [], []
    for v, d in vpicks:
        base = r.coefs[v, d]
        r.coefs[v, d] = base + h
        jp = r.adjoint(cfg)[0]
        r.coefs[v, d] = base - h
        jm = r.adjoint(cfg)[0]
        r.coefs[v, d] = base
        fd_a.append((jp - jm) / (2 * h))
        ad_a.append(dA[v, d])
    fd_g, ad_g, fd_a, ad_a = map(np.array, (fd_g, ad_g, fd_a, ad_a))
    assert np.max(np.abs(fd_g - ad_g)) <= 1e-6 * np.max(np.abs(fd_g)), (fd_g, ad_g)
    assert np.max(np.abs(fd_a - ad_a)) <= 1e-6 * np.max(np.abs(fd_a)), (fd_a, ad_a)
    # missing connections have no gradient
    assert np.all(dG[p.output_state_indices < 0] == 0)
    assert np.all(dA[p.geometry.neighbours < 0] == 0)


def test_adjoint_objective_matches_epoch_losses():
    # J is the sum over steps of (s_mse/2)|r|^2 + (s_reg/2) sum f^2, i.e.
    # T/2 * (loss_mse_batch + loss of the max-reg part per step), recomputed
    # here from the epoch's per-step arrays of the same parameters.
    p = problem((3, 3, 3), 40, 6, 4)
    cfg = Algorithm(learning_rate=0.0, maximum_regularization_threshold=0.05, freeze_delays=False)
    r = oracle.OracleResults(p, np.float64)
    J, _, _ = r.adjoint(cfg)
    r.run_epoch(cfg)
    mse_part = 0.5 * np.sum(r.loss_mse)  # loss_mse = |r|^2 / S, s_mse = 1/S
    mr = np.diff(np.concatenate([[0.0], r.loss_maximum_regularization]))
    assert J == pytest.approx(mse_part + 0.5 * np.sum(mr), rel=1e-12)


def test_f32_adjoint_close_to_f64():
    p = problem((3, 3, 3), 40, 6, 5)
    cfg = Algorithm(maximum_regularization_threshold=0.05)
    J64, dG64, dA64 = oracle.OracleResults(p, np.float64).adjoint(cfg)
    J32, dG32, dA32 = oracle.OracleResults(p, np.float32).adjoint(cfg)
    assert J32 == pytest.approx(J64, rel=1e-5)
    assert np.max(np.abs(dG32 - dG64)) <= 1e-4 * np.max(np.abs(dG64))
    assert np.max(np.abs(dA32 - dA64)) <= 1e-4 * np.max(np.abs(dA64))
