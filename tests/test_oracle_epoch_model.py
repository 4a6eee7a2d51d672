"""The oracle's multi-beat epoch, Adam update and smoothness term against an
independent float64 numpy model (tests/numpy_model.py: EpochModel), and its
FP32 Adam step against the update formula restated in numpy float32.

What these pin (reference citations relative to src/core/):
  * run_epoch's beat loop (algorithm.rs:124-281): explicit beat order, update
    batches of batch_size beats with a partial last batch, lr / (T *
    batch_size) (update.rs:35-38), derivatives / FIR / IIR / running
    maximum-regularisation sum reset only at the epoch start and after a full
    batch (derivation.rs:114-122), per-beat projection with the beat's own
    measurement matrix, batch losses from the last-visited beat
    (metrics.rs:204-219);
  * Adam with bias correction at the running step (update.rs:119-146,
    173-200), over several updates and epochs;
  * the smoothness derivative with voxels whose average delay is None
    (derivation.rs:241-302, 554-594).
A plausible slip in any of them (resetting FIR/IIR per beat, dividing by
T * n for a partial batch, a step counter from 0, a missing own term in the
neighbourhood mean, counting an invalid neighbour) changes an array the
comparison covers by far more than its tolerance.
"""
import numpy as np
import pytest

from cardiotrust_amd import Algorithm, APDerivative, Optimizer, synth
from oracle import oracle

from numpy_model import EpochModel


def multibeat_problem(beats=3, shape=(3, 3, 2), T=36, max_delay=5, seed=21, zero_voxel=None):
    g = synth.grid_geometry(np.ones(shape, dtype=bool), 2.5)
    offsets = synth.grid_motion_offsets((beats, 1, 1), (10.0 * (beats - 1), 0.0, 0.0))
    p = synth.make_problem(g, num_steps=T, max_delay=max_delay, seed=seed, num_beats=beats,
                           beat_offsets_mm=offsets, gain_sigma=0.1)
    # drive every state so that all connections carry signal
    p.control_matrix[:] = (0.1 * (0.5 + synth.uniform(seed, 77, p.num_states))).astype(np.float32)
    # coefficients a hair inside both margins: an update moves some across,
    # which rolls their delays (update.rs:206-228)
    p.coefs[1, :] = np.float32(1.00001e-4)
    p.coefs[2, :] = np.float32(1.0 - 1.00001e-4)
    if zero_voxel is not None:
        # every gain of one voxel zero: its average delay is None
        p.gains[3 * zero_voxel:3 * zero_voxel + 3, :] = 0.0
    p.measurements = np.ascontiguousarray(oracle.simulate_f32(p.truth()), dtype=np.float32)
    assert p.measurement_matrix.shape[0] == beats
    assert not np.array_equal(p.measurement_matrix[0], p.measurement_matrix[-1])
    return p


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    scale = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / scale) if scale > 0 else float(np.max(np.abs(a), initial=0))


def expand(y):
    """(N, 26, 3) per-connection values as the reference's (3N, 78) arrays."""
    N = y.shape[0]
    return np.repeat(y.reshape(N, 78), 3, axis=0)


CASES = [
    dict(name="sgd-all-beats", beats=3, epochs=2, order=[2, 0, 1],
         cfg=dict(learning_rate=0.5, batch_size=0, freeze_delays=False)),
    dict(name="sgd-batch2-partial", beats=3, epochs=2, order=[1, 2, 0],
         cfg=dict(learning_rate=0.5, batch_size=2, freeze_delays=False)),
    dict(name="adam-batch2", beats=3, epochs=3, order=[2, 0, 1], optimizer=1,
         cfg=dict(learning_rate=0.01, batch_size=2, optimizer=Optimizer.Adam,
                  freeze_delays=False)),
    dict(name="adam-all-beats", beats=2, epochs=3, order=[1, 0], optimizer=1,
         cfg=dict(learning_rate=0.01, batch_size=0, optimizer=Optimizer.Adam,
                  freeze_delays=False)),
    dict(name="smoothness-none-voxel-batch1", beats=2, epochs=2, order=[0, 1], zero_voxel=4,
         cfg=dict(learning_rate=0.2, batch_size=1, freeze_delays=False, freeze_gains=True,
                  smoothness_regularization_strength=0.1)),
    dict(name="smoothness-none-voxel-all", beats=3, epochs=1, order=[2, 1, 0], zero_voxel=7,
         cfg=dict(learning_rate=0.2, batch_size=0, freeze_delays=False,
                  smoothness_regularization_strength=0.25)),
    dict(name="simple-diffreg-batch2", beats=3, epochs=2, order=[0, 2, 1],
         cfg=dict(learning_rate=0.3, batch_size=2, freeze_delays=False,
                  ap_derivative=APDerivative.Simple, difference_regularization_strength=1e-2)),
    dict(name="frozen-gains-adam", beats=2, epochs=2, order=[1, 0], optimizer=1,
         cfg=dict(learning_rate=0.01, batch_size=1, optimizer=Optimizer.Adam,
                  freeze_gains=True, freeze_delays=False)),
]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_multibeat_epochs_match_numpy_model(case):
    p = multibeat_problem(beats=case["beats"], zero_voxel=case.get("zero_voxel"))
    cfg = Algorithm(maximum_regularization_threshold=0.05, **case["cfg"])
    opt = case.get("optimizer", 0)
    epochs = case["epochs"]
    per_epoch = 1 if not cfg.batch_size else -(-p.num_beats // cfg.batch_size)
    o = oracle.OracleResults(p, np.float64, optimizer=opt, num_batch_entries=epochs * per_epoch)
    m = EpochModel(p, optimizer=opt)
    for _ in range(epochs):
        o.run_epoch(cfg, case["order"])
        m.run_epoch(cfg, case["order"])
    assert o.batch_index.value == m.batch_index == epochs * per_epoch
    assert o.d_step[0] == m.step == 1 + epochs * per_epoch
    assert np.array_equal(o.delays, m.delays)
    if not cfg.freeze_delays:
        assert np.any(o.delays != p.delays)  # some delays rolled
    tol = 1e-10
    checks = {
        "gains": (o.gains, m.gains_rows), "coefs": (o.coefs, m.coefs),
        "system_states": (o.system_states, m.system_states),
        "ap_outputs_now": (o.ap_outputs_now, expand(m.filter_outputs)),
        "d_gains": (o.d_gains, m.d_gains), "d_coefs": (o.d_coefs, m.d_coefs),
        "d_coefs_fir": (o.d_coefs_fir, expand(m.fir)), "d_coefs_iir": (o.d_coefs_iir, expand(m.iir)),
        "maximum_regularization_sum": (o.maximum_regularization_sum, [m.mr_sum]),
        "loss": (o.loss, m.loss), "loss_mse": (o.loss_mse, m.loss_mse),
        "loss_maximum_regularization": (o.loss_maximum_regularization, m.loss_mr),
        "loss_batch": (o.loss_batch, m.loss_batch), "loss_mse_batch": (o.loss_mse_batch, m.loss_mse_batch),
        "loss_maximum_regularization_batch": (o.loss_maximum_regularization_batch, m.loss_mr_batch),
    }
    if cfg.smoothness_regularization_strength:
        avg, valid = m.avg, m.avg_valid  # as of the last update
        assert np.array_equal(o.average_delays_valid.astype(bool), valid)
        assert not valid.all()
        checks["average_delays"] = (o.average_delays[valid], avg[valid])
    if opt:
        checks["gains_first_moment"] = (o.gains_first_moment, m.moments["gains"].reshape(-1, 78))
        checks["coefs_second_moment"] = (o.coefs_second_moment, m.moments2["coefs"])
    for k, (got, want) in checks.items():
        assert rel(got, want) < tol, (k, rel(got, want))


def test_model_distinguishes_per_beat_filter_reset():
    """The pins have teeth: a model that restarts FIR / IIR every beat gives
    coefficient derivatives the oracle's do not match."""
    p = multibeat_problem(beats=3)
    cfg = Algorithm(learning_rate=0.5, batch_size=0, freeze_delays=False)
    o = oracle.OracleResults(p, np.float64)
    o.run_epoch(cfg, [0, 1, 2])
    m = EpochModel(p)
    beat = m._beat

    def reset_each_beat(b, c, carry):
        carry["fir"][:] = 0.0
        carry["iir"][:] = 0.0
        beat(b, c, carry)

    m._beat = reset_each_beat
    m.run_epoch(cfg, [0, 1, 2])
    assert rel(o.d_coefs, m.d_coefs) > 1e-3


def f32_adam(v, d, m1, m2, step, lr, bs):
    """update.rs:119-146 in float32 numpy (one f32 rounding per operation):
    m1 = m1 b1 + (1 - b1) d; m2 = m2 b2 + ((1 - b2) d) d;
    v -= (lr / bs) * (m1 / (1 - b1^k)) / (sqrt(m2 / (1 - b2^k)) + eps)."""
    F = np.float32
    b1, b2, eps = F(0.9), F(0.999), F(1e-8)
    m1 = (m1 * b1 + (F(1) - b1) * d).astype(F)
    m2 = (m2 * b2 + (F(1) - b2) * d * d).astype(F)
    c1 = F(1) - np.power(b1, F(step), dtype=F)
    c2 = F(1) - np.power(b2, F(step), dtype=F)
    factor = (m1 / c1) / (np.sqrt(m2 / c2) + eps)
    return (v - (F(lr) / F(bs)) * factor).astype(F), m1, m2


def f32_roll(a, D):
    """update.rs:206-228 in float32."""
    F = np.float32
    mg = F(1e-4)
    hi, lo = a > F(1) - mg, a < mg
    a2 = np.where(hi, np.where(D > 1, F(2) * mg, F(1) - mg),
                  np.where(lo, np.where(D < 1000, F(1) - F(2) * mg, mg), a)).astype(F)
    D2 = np.where(hi & (D > 1), D - 1, np.where(lo & (D < 1000), D + 1, D)).astype(np.int32)
    return a2, D2


def test_oracle_adam_f32_bit_exact_over_steps():
    """Three epochs of one update each: the FP32 oracle's parameters and
    moments after update k equal the float32 restatement of the reference's
    Adam at step k, applied to the oracle's own derivatives of that epoch, bit
    for bit (steps 1, 2, 3: the bias corrections differ at every step)."""
    p = multibeat_problem(beats=2)
    cfg = Algorithm(learning_rate=0.01, batch_size=0, optimizer=Optimizer.Adam,
                    freeze_delays=False)
    o = oracle.OracleResults(p, np.float32, optimizer=1, num_batch_entries=3)
    bs = p.num_steps * p.num_beats
    for k in (1, 2, 3):
        before = {n: getattr(o, n).copy() for n in
                  ("gains", "coefs", "delays", "gains_first_moment", "gains_second_moment",
                   "coefs_first_moment", "coefs_second_moment")}
        assert o.d_step[0] == k
        o.run_epoch(cfg, [1, 0])
        g, gm1, gm2 = f32_adam(before["gains"], o.d_gains, before["gains_first_moment"],
                               before["gains_second_moment"], k, cfg.learning_rate, bs)
        a, am1, am2 = f32_adam(before["coefs"], o.d_coefs, before["coefs_first_moment"],
                               before["coefs_second_moment"], k, cfg.learning_rate, bs)
        a, D = f32_roll(a, before["delays"])
        for name, want in (("gains", g), ("gains_first_moment", gm1), ("gains_second_moment", gm2),
                           ("coefs", a), ("coefs_first_moment", am1),
                           ("coefs_second_moment", am2)):
            got = getattr(o, name)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (k, name)
        assert np.array_equal(o.delays, D)
        assert np.any(g != before["gains"])
