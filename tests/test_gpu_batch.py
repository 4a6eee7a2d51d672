"""Batched epochs (ct_batch_*) against each scenario's own epoch.

The batch runs every member's epoch with the member's own order of
operations, only laid out differently on the device (rounds of beats, one
persistent propagation launch per round, optional graph replay). The bar is
therefore bit-exact equality of every result array with the same scenario
run alone through ct_epoch_kernel_execute, which the parity tests pin to the
CPU oracle; one member is also compared with the oracle directly.
"""
import sys

import numpy as np
import pytest

from cardiotrust_amd import Algorithm, APDerivative, Optimizer, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

FIELDS = ("gains", "coefs", "delays", "ap_outputs_now", "ap_outputs_last", "system_states",
          "measurements", "residuals", "d_gains", "d_coefs", "d_coefs_iir", "d_coefs_fir",
          "mapped_residuals", "maximum_regularization", "maximum_regularization_sum",
          "average_delays", "loss", "loss_mse", "loss_maximum_regularization", "loss_batch",
          "loss_mse_batch", "loss_maximum_regularization_batch", "derivatives_step")


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    from cardiotrust_amd.gpu import GPU
    g = GPU(0)
    yield g
    g.close()


def problems():
    """Uneven scenarios: sizes 40-1100 voxels, 1-3 beats, 30-80 steps."""
    specs = [(1100, 45, 3, 12), (300, 60, 2, 20), (40, 80, 1, 9), (216, 30, 2, 6),
             (700, 52, 1, 15)]
    out = []
    for i, (n, T, B, maxd) in enumerate(specs):
        g = synth.blob(100 + i, n) if n != 216 else synth.cube(6)
        offsets = synth.grid_motion_offsets((B, 1, 1), (10.0 * (B - 1), 0.0, 0.0))
        out.append(synth.make_problem(g, num_steps=T, max_delay=maxd, seed=100 + i, num_beats=B,
                                      beat_offsets_mm=offsets, gain_sigma=0.2,
                                      simulate=oracle.simulate_f32))
    return out


CFGS = [
    dict(cfg=Algorithm(learning_rate=0.05, freeze_delays=False), optimizer=0),
    dict(cfg=Algorithm(learning_rate=0.01, batch_size=1, optimizer=Optimizer.Adam,
                       freeze_delays=False), optimizer=1),
    dict(cfg=Algorithm(learning_rate=0.05, freeze_delays=False,
                       ap_derivative=APDerivative.Simple, maximum_regularization_threshold=0.05),
         optimizer=0),
    dict(cfg=Algorithm(learning_rate=0.02, freeze_delays=False,
                       difference_regularization_strength=0.1), optimizer=0),
    dict(cfg=Algorithm(learning_rate=0.05, freeze_gains=True, freeze_delays=False), optimizer=0),
]


class Member:
    def __init__(self, gpu, p, spec, epochs, order=None, gradient=None, mode=0, share=None):
        from cardiotrust_amd.gpu import Buffer, EpochKernel, ResultsGPU
        nbatch = epochs * (p.num_beats if spec["cfg"].batch_size else 1)
        self.r = ResultsGPU(gpu, p, num_batch_entries=nbatch, optimizer=spec["optimizer"])
        self.r.set_propagation(mode)
        self.buf = Buffer(gpu, p.measurements)
        self.k = EpochKernel(gpu, self.r, self.buf, spec["cfg"], 3 * p.num_voxels, p.num_sensors,
                             p.num_steps)
        if order is not None:
            self.k.set_beat_order(order)
        if gradient is not None:
            self.k.set_gradient(*gradient)
        if share is not None:
            self.k.set_share_propagation(share)

    def results(self):
        out = self.r.update_from_gpu(list(FIELDS))
        out["batch_index"] = np.array([self.k.batch_index()])
        return out

    def close(self):
        self.k.close()
        self.buf.close()
        self.r.close()


def propagations(p, spec):
    """Propagation jobs of one epoch under the default share mode: one per
    update batch of beats (the reference derivation shares the identical
    propagation of the beats of one update batch)."""
    group = spec["cfg"].batch_size or p.num_beats
    return -(-p.num_beats // group) if p.num_beats else 0


def orders(ps):
    return [list(range(p.num_beats))[::-1] if p.num_beats > 1 else None for p in ps]


def run_alone(gpu, ps, epochs, gradients=None, modes=None, shares=None):
    outs = []
    for i, p in enumerate(ps):
        m = Member(gpu, p, CFGS[i % len(CFGS)], epochs, orders(ps)[i],
                   gradients[i] if gradients else None, modes[i] if modes else 0,
                   shares[i] if shares else None)
        for _ in range(epochs):
            m.k.execute()
        outs.append(m.results())
        m.close()
    return outs


def run_batched(gpu, ps, epochs, vpg=0, graph=False, gradients=None, modes=None, shares=None):
    from cardiotrust_amd.gpu import Batch
    ms = [Member(gpu, p, CFGS[i % len(CFGS)], epochs, orders(ps)[i],
                 gradients[i] if gradients else None, modes[i] if modes else 0,
                 shares[i] if shares else None)
          for i, p in enumerate(ps)]
    b = Batch(gpu, [m.k for m in ms], voxels_per_group=vpg, graph=graph)
    infos = []
    for _ in range(epochs):
        b.execute()
        infos.append(b.info())
    outs = [m.results() for m in ms]
    b.close()
    for m in ms:
        m.close()
    return outs, infos


def assert_same(alone, batched):
    for s, (a, b) in enumerate(zip(alone, batched)):
        for k in list(FIELDS) + ["batch_index"]:
            x, y = np.asarray(a[k]), np.asarray(b[k])
            assert x.shape == y.shape, (s, k)
            if x.dtype.kind == "f":
                same = np.array_equal(x.view(np.uint32 if x.itemsize == 4 else np.uint64),
                                      y.view(np.uint32 if y.itemsize == 4 else np.uint64))
            else:
                same = np.array_equal(x, y)
            assert same, f"scenario {s}: {k} differs (max |d| "\
                         f"{np.nanmax(np.abs(x.astype(np.float64) - y.astype(np.float64)))})"


@pytest.fixture(scope="module")
def ps():
    return problems()


@pytest.fixture(scope="module")
def alone2(gpu, ps):
    return run_alone(gpu, ps, epochs=2)


@pytest.mark.parametrize("vpg", [472, 48, 0])
def test_batch_equals_each_scenario_alone(gpu, ps, alone2, vpg):
    note(f"case: batch of {len(ps)} uneven scenarios, 2 epochs, voxels per group {vpg}")
    got, infos = run_batched(gpu, ps, epochs=2, vpg=vpg)
    assert_same(alone2, got)
    info = infos[-1]
    assert info["side_streams"] == 3  # reference derivation only: beats overlap on side streams
    assert info["rounds"] == max(p.num_beats for p in ps)
    assert info["jobs"] == sum(propagations(p, CFGS[i % len(CFGS)]) for i, p in enumerate(ps))
    assert info["slots"] > 0
    if vpg == 0:
        assert info["autotune"], info
        assert info["voxels_per_group"] in info["autotune"]
    else:
        assert info["voxels_per_group"] == vpg


def test_batch_graph_replays_equal_eager_epochs(gpu, ps):
    note("case: batch captured in a graph, 4 epochs (1 eager + 3 replays) vs alone")
    alone = run_alone(gpu, ps, epochs=4)
    got, infos = run_batched(gpu, ps, epochs=4, vpg=64, graph=True)
    assert [i["graph"] for i in infos] == [0, 1, 1, 1]
    assert_same(alone, got)


def test_batch_adjoint_members_and_graph(gpu, ps):
    note("case: batch with reverse-mode members (K=16, K=7), eager and graph, 3 epochs")
    grads = [(1, 16), None, (1, 7), None, (1, 16)]
    alone = run_alone(gpu, ps, epochs=3, gradients=grads)
    got, _ = run_batched(gpu, ps, epochs=3, vpg=96, gradients=grads)
    assert_same(alone, got)
    got_g, infos = run_batched(gpu, ps, epochs=3, vpg=96, graph=True, gradients=grads)
    assert infos[-1]["graph"] == 1
    # the reverse-mode members share one backward pass per update batch, so
    # their sweeps wait for the batched sweep launch and the beats' other
    # work runs on the side streams
    assert infos[-1]["side_streams"] == 3
    # the three members' sweeps (one update batch each) in the batched sweep launches
    assert infos[-1]["sweep_jobs"] == 3
    assert_same(alone, got_g)


def test_batch_all_reverse_mode_sweeps_per_round(gpu, ps):
    """Every member in reverse mode. A member's backward pass (and its update)
    waits until the member propagates again or the epoch ends, and the pending
    passes sweep together in one batched launch. Scenario 1 updates after each
    beat (batch_size 1, Adam): its first pass is flushed before round 1's
    propagation together with those of scenarios 2 and 4 (one beat each), its
    second at the end with those of scenarios 0 and 3. Eager and
    graph-replayed epochs against each scenario alone, bit for bit."""
    note("case: batch of 5, all reverse mode K=16, sweeps batched per round, eager and graph")
    grads = [(1, 16)] * len(ps)
    alone = run_alone(gpu, ps, epochs=3, gradients=grads)
    got, infos = run_batched(gpu, ps, epochs=3, vpg=128, gradients=grads)
    groups = sum(propagations(p, CFGS[i % len(CFGS)]) for i, p in enumerate(ps))
    assert infos[-1]["sweep_jobs"] == groups == 6, infos[-1]
    assert_same(alone, got)
    got_g, infos = run_batched(gpu, ps, epochs=3, vpg=128, graph=True, gradients=grads)
    assert infos[-1]["graph"] == 1 and infos[-1]["sweep_jobs"] == groups
    assert_same(alone, got_g)


def test_batch_graph_recaptures_past_16_reverse_mode_epochs(gpu, ps):
    """Reverse-mode members in a graph-replayed batch: the captured epoch's
    q rings hold room for 16 epochs of delay growth, so the batch captures
    again after 16 replays. 18 epochs (one eager, replays across the
    recapture), the backward passes in the batched sweep launches, against
    each scenario alone, bit for bit."""
    note("case: batch of 5, all reverse mode K=16, graph, 18 epochs across the recapture")
    grads = [(1, 16)] * len(ps)
    alone = run_alone(gpu, ps, epochs=18, gradients=grads)
    got, infos = run_batched(gpu, ps, epochs=18, vpg=128, graph=True, gradients=grads)
    assert [i["graph"] for i in infos] == [0] + [1] * 17
    assert infos[-1]["sweep_jobs"] == 6
    assert_same(alone, got)


def test_batch_with_a_member_sweeping_inside_its_beats(gpu, ps):
    """Scenario 0 in reverse mode with one backward pass per beat (shared
    propagation off): its sweeps run inside its beats, so the batch keeps
    every scenario on the GPU's stream; the other reverse-mode member still
    sweeps in the batched launch. Against each scenario alone, bit for bit."""
    note("case: batch with a reverse-mode member that does not share (sweeps in its beats)")
    grads = [(1, 16), None, (1, 16), None, None]
    shares = [False, None, None, None, None]
    alone = run_alone(gpu, ps, epochs=2, gradients=grads, shares=shares)
    got, infos = run_batched(gpu, ps, epochs=2, vpg=96, gradients=grads, shares=shares)
    assert infos[-1]["side_streams"] == 0
    assert infos[-1]["sweep_jobs"] == 1
    assert_same(alone, got)


def test_batch_with_a_per_step_member(gpu, ps):
    note("case: batch where scenario 1 uses the per-step propagation")
    modes = [0, 1, 0, 0, 0]
    alone = run_alone(gpu, ps, epochs=2, modes=modes)
    got, infos = run_batched(gpu, ps, epochs=2, vpg=128, modes=modes)
    assert_same(alone, got)
    assert infos[-1]["jobs"] == sum(propagations(p, CFGS[i % len(CFGS)])
                                   for i, p in enumerate(ps) if modes[i] != 1)


def test_batch_member_matches_oracle(gpu, ps):
    note("case: batch member 0 vs the FP32 oracle, 1 epoch")
    got, _ = run_batched(gpu, ps, epochs=1, vpg=64)
    p = ps[0]
    spec = CFGS[0]
    o = oracle.OracleResults(p, np.float32, optimizer=spec["optimizer"], num_batch_entries=1)
    o.run_epoch(spec["cfg"], orders(ps)[0])
    for k in ("system_states", "ap_outputs_now", "ap_outputs_last", "delays"):
        assert np.array_equal(got[0][k], getattr(o, k)), k
    for k in ("gains", "coefs", "d_gains", "d_coefs", "loss_batch"):
        a, b = got[0][k].astype(np.float64), getattr(o, k).astype(np.float64)
        assert np.max(np.abs(a - b)) <= 1e-4 * max(np.max(np.abs(b)), 1e-30), k


def test_batch_rejects_bad_members(gpu, ps):
    from cardiotrust_amd.gpu import Batch, CTError
    note("case: batch argument checks")
    m = Member(gpu, ps[2], CFGS[0], 1)
    with pytest.raises(CTError):
        Batch(gpu, [])
    with pytest.raises(CTError):
        Batch(gpu, [m.k, m.k])
    b = Batch(gpu, [m.k])
    with pytest.raises(CTError):
        b.set_voxels_per_group(1000)
    b.close()
    m.close()


def test_batch_with_an_empty_scenario(gpu, ps):
    note("case: batch with an empty scenario (N = 0) among others")
    g = synth.grid_geometry(np.zeros((2, 2, 2), dtype=bool), 2.5)
    empty = synth.make_problem(g, num_steps=30, max_delay=4, seed=1, simulate=oracle.simulate_f32)
    mixed = [ps[2], empty, ps[3]]
    alone = run_alone(gpu, mixed, epochs=2)
    got, infos = run_batched(gpu, mixed, epochs=2, vpg=64)
    assert_same(alone, got)
    assert infos[-1]["scenarios"] == 3


def test_batch_wait_timeout_leaves_parameters(gpu, ps):
    """A failure forced from the host through the kernels' own bounded waits:
    a wait limit of 0 s on a reverse-mode member bounds that member's backward
    sweep and, as the shortest limit of the batch, every wait of the batched
    propagation launch. A wait that is not satisfied at its first poll then
    raises the abort word; every workgroup leaves at its next wait (so every
    kernel ends normally), the members' updates read the word and skip, and
    ct_batch_execute reports CT_ERR_DEVICE with every member's parameters and
    epoch counters unchanged. With the limit restored the batch runs on the
    paths that did not time out, and two more epochs equal two epochs of each
    scenario alone bit for bit. (An attempt in which no wait has to wait
    completes normally; one of three must time out.)"""
    from cardiotrust_amd.gpu import Batch, CTError
    note("case: batch of 5 with reverse-mode members, member 0 wait limit 0 s, then 2 epochs")
    grads = [(1, 16), None, (1, 7), None, (1, 16)]
    capacity = 5
    ms = [Member(gpu, p, CFGS[i % len(CFGS)], capacity, orders(ps)[i], grads[i])
          for i, p in enumerate(ps)]
    b = Batch(gpu, [m.k for m in ms], voxels_per_group=96)
    ms[0].r.set_wait_limit(0.0)
    before = [m.results() for m in ms]
    ok_epochs, failed = 0, False
    for _ in range(3):
        try:
            b.execute()
        except CTError as e:
            assert "timed out" in str(e), e
            failed = True
            break
        ok_epochs += 1
        before = [m.results() for m in ms]
    assert failed
    after = [m.results() for m in ms]
    for s, (x, y) in enumerate(zip(before, after)):
        for k in ("gains", "coefs", "delays", "batch_index", "derivatives_step"):
            assert np.array_equal(x[k], y[k]), (s, k)
    ms[0].r.set_wait_limit(2.0)
    for _ in range(2):
        b.execute()
    got = [m.results() for m in ms]
    b.close()
    for m in ms:
        m.close()
    alone = []
    for i, p in enumerate(ps):
        m = Member(gpu, p, CFGS[i % len(CFGS)], capacity, orders(ps)[i], grads[i])
        for _ in range(ok_epochs + 2):
            m.k.execute()
        alone.append(m.results())
        m.close()
    assert_same(alone, got)
