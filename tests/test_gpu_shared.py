"""Shared propagation (ct_epoch_kernel_set_share_propagation).

The beats of one update batch have the same parameters and drive, and the
propagation does not depend on the beat (prediction.rs:15-105), so running it
once per update batch must leave every result bit-identical under the
reference derivation. Under the reverse-mode gradient the update batch's
summed objective sum_b J_b is differentiated with one backward pass: checked
against the oracle's FP64 gradient summed over the beats (itself pinned by
central differences) to 1e-3 relative, and against the per-beat passes.
"""
import sys

import numpy as np
import pytest

from cardiotrust_amd import Algorithm, Optimizer, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

FIELDS = ("gains", "coefs", "delays", "ap_outputs_now", "ap_outputs_last", "system_states",
          "measurements", "d_gains", "d_coefs", "d_coefs_iir", "d_coefs_fir", "loss",
          "loss_mse", "loss_maximum_regularization", "loss_batch", "loss_mse_batch",
          "loss_maximum_regularization_batch", "maximum_regularization_sum", "average_delays")


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


def problem(seed, beats=3, T=60, n=400):
    g = synth.blob(200 + seed, n)
    offsets = synth.grid_motion_offsets((beats, 1, 1), (10.0 * (beats - 1), 0.0, 0.0))
    return synth.make_problem(g, num_steps=T, max_delay=12, seed=200 + seed, num_beats=beats,
                              beat_offsets_mm=offsets, gain_sigma=0.25,
                              simulate=oracle.simulate_f32)


def run(gpu, p, cfg, share, epochs=2, gradient=None, optimizer=0, order=None, batch=False,
        graph=False, launches=None):
    from cardiotrust_amd.gpu import Batch, Buffer, EpochKernel, ResultsGPU
    nb = epochs * (p.num_beats if cfg.batch_size else 1)
    r = ResultsGPU(gpu, p, num_batch_entries=nb, optimizer=optimizer)
    buf = Buffer(gpu, p.measurements)
    k = EpochKernel(gpu, r, buf, cfg, 3 * p.num_voxels, p.num_sensors, p.num_steps)
    if gradient:
        k.set_gradient(*gradient)
    if order is not None:
        k.set_beat_order(order)
    if share != "default":
        k.set_share_propagation(share)
    k.set_profiling(launches is not None)
    b = Batch(gpu, [k], voxels_per_group=96, graph=graph) if batch else None
    for _ in range(epochs):
        (b or k).execute()
        if launches is not None:
            launches.append(int(k.profile()["propagate_launches"]))
    if b:
        b.close()
    out = r.update_from_gpu(list(FIELDS))
    k.close()
    buf.close()
    r.close()
    return out


@pytest.mark.parametrize("batch_size,optimizer", [(0, 0), (2, 1)])
def test_shared_propagation_is_bit_identical_under_the_reference_derivation(gpu, batch_size,
                                                                           optimizer):
    note(f"case: 3 beats, batch_size {batch_size}, optimizer {optimizer}, share on vs off")
    p = problem(1)
    cfg = Algorithm(learning_rate=0.02, batch_size=batch_size,
                    optimizer=Optimizer.Adam if optimizer else Optimizer.Sgd,
                    freeze_delays=False)
    a = run(gpu, p, cfg, False, optimizer=optimizer, order=[1, 2, 0])
    b = run(gpu, p, cfg, True, optimizer=optimizer, order=[1, 2, 0])
    for f in FIELDS:
        x, y = np.asarray(a[f]), np.asarray(b[f])
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f


def test_shared_reverse_mode_is_the_gradient_of_the_summed_objective(gpu):
    note("case: 3 beats, gradient mode 1 with one backward pass per update batch")
    p = problem(2)
    cfg = Algorithm(learning_rate=0.0, freeze_delays=False, maximum_regularization_threshold=0.1)
    got = run(gpu, p, cfg, True, epochs=1, gradient=(1, 16))
    per_beat = run(gpu, p, cfg, False, epochs=1, gradient=(1, 16))
    o = oracle.OracleResults(p, np.float64)
    dG = sum(o.adjoint(cfg, beat=b)[1] for b in range(p.num_beats))
    dA = sum(o.adjoint(cfg, beat=b)[2] for b in range(p.num_beats))
    eg, ea = rel_inf(got["d_gains"], dG), rel_inf(got["d_coefs"], dA)
    note(f"  vs FP64 sum over beats: dG {eg:.2e}, dA {ea:.2e}; vs per-beat passes: "
         f"{rel_inf(got['d_gains'], per_beat['d_gains']):.2e}")
    assert eg <= 1e-3 and ea <= 1e-3
    assert rel_inf(got["d_gains"], per_beat["d_gains"]) <= 1e-4
    assert rel_inf(got["d_coefs"], per_beat["d_coefs"]) <= 1e-4


def test_shared_propagation_in_a_batch_equals_alone(gpu):
    note("case: shared propagation, reverse mode, through ct_batch vs alone")
    p = problem(3, beats=4)
    cfg = Algorithm(learning_rate=0.05, batch_size=2, freeze_delays=False)
    a = run(gpu, p, cfg, True, gradient=(1, 16))
    b = run(gpu, p, cfg, True, gradient=(1, 16), batch=True)
    for f in FIELDS:
        x, y = np.asarray(a[f]), np.asarray(b[f])
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f


def test_shared_propagation_in_a_graph_replayed_batch(gpu):
    note("case: shared propagation, reverse mode, batch captured in a graph, 3 epochs")
    p = problem(4, beats=3)
    cfg = Algorithm(learning_rate=0.05, freeze_delays=False)
    a = run(gpu, p, cfg, True, epochs=3, gradient=(1, 16))
    b = run(gpu, p, cfg, True, epochs=3, gradient=(1, 16), batch=True, graph=True)
    for f in FIELDS:
        x, y = np.asarray(a[f]), np.asarray(b[f])
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), f



def test_default_shares_under_both_gradients(gpu):
    """The default (automatic) propagates once per update batch under both
    gradients: under the reference derivation it is bit-identical to one
    propagation per beat; under the reverse-mode gradient it is the explicit
    shared mode, one forward and one backward pass per update batch."""
    note("case: 3 beats, default share mode vs explicit, both gradient modes")
    p = problem(5)
    cfg = Algorithm(learning_rate=0.02, freeze_delays=False)
    n_auto, n_off = [], []
    auto = run(gpu, p, cfg, "default", launches=n_auto)
    off = run(gpu, p, cfg, False, launches=n_off)
    assert n_auto == [1, 1] and n_off == [3, 3], (n_auto, n_off)
    for f in FIELDS:
        assert np.array_equal(np.asarray(auto[f]).view(np.uint8), np.asarray(off[f]).view(np.uint8)), f
    n_auto = []
    auto = run(gpu, p, cfg, "default", gradient=(1, 16), launches=n_auto)
    on = run(gpu, p, cfg, True, gradient=(1, 16))
    assert n_auto == [1, 1], n_auto
    for f in FIELDS:
        assert np.array_equal(np.asarray(auto[f]).view(np.uint8), np.asarray(on[f]).view(np.uint8)), f


def test_shared_reverse_mode_over_more_beats_than_one_forcing_pass(gpu):
    """Ten beats in one update batch: the forcing of the shared backward pass
    sums up to eight beats per launch, so it takes two launches, the second
    adding to the first's output. Against the per-beat passes and the FP64
    oracle summed over the beats, with an explicit beat order."""
    note("case: 10 beats, batch_size 0, reverse mode shared vs per beat, beat order reversed")
    p = problem(6, beats=10, T=48, n=300)
    cfg = Algorithm(learning_rate=0.0, freeze_delays=False, maximum_regularization_threshold=0.1)
    order = list(range(p.num_beats))[::-1]
    got = run(gpu, p, cfg, True, epochs=1, gradient=(1, 16), order=order)
    per_beat = run(gpu, p, cfg, False, epochs=1, gradient=(1, 16), order=order)
    o = oracle.OracleResults(p, np.float64)
    dG = sum(o.adjoint(cfg, beat=b)[1] for b in range(p.num_beats))
    dA = sum(o.adjoint(cfg, beat=b)[2] for b in range(p.num_beats))
    eg, ea = rel_inf(got["d_gains"], dG), rel_inf(got["d_coefs"], dA)
    note(f"  vs FP64 sum over beats: dG {eg:.2e}, dA {ea:.2e}; vs per-beat passes: "
         f"{rel_inf(got['d_gains'], per_beat['d_gains']):.2e}")
    assert eg <= 1e-3 and ea <= 1e-3
    assert rel_inf(got["d_gains"], per_beat["d_gains"]) <= 1e-4
    assert rel_inf(got["d_coefs"], per_beat["d_coefs"]) <= 1e-4
