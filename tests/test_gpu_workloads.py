"""The batched workloads at their stated sizes.

Uneven workload: 64 seeded blob scenarios (voxels log-uniform in 1k-40k,
1-8 beats, 600-2000 steps, delays 1-20), the reference derivation, default
shared propagation and side streams, graph replay. Over two epochs every
result array of every scenario must equal the same scenario run alone
through ct_epoch_kernel_execute, bit for bit (compared by SHA-256, so the host
holds one member's arrays at a time), and the cheapest scenario's first epoch
alone is compared with the FP32 oracle (states and delays bit-exact, the rest
at 1e-4).

The same 64 scenarios in reverse mode (checkpoints every 16 steps): every
scenario's backward pass is left pending by its last beat and all 64 sweeps
run as jobs of one batched sweep launch at the end of the epoch; two epochs
(one eager, one graph replay) must equal each scenario alone bit for bit.

Shell workload: 8 ellipsoidal shells of about 120k voxels with slow
sub-regions, 4 beats, T = 2000, the reverse-mode gradient with checkpoints
every 64 steps, one propagation and one backward pass per epoch (the default
shared mode). Two epochs (one eager, one graph replay) must equal each scenario
alone bit for bit. Member 0 runs with learning rate 0, so its derivative
arrays hold the gradient of its summed objective sum_b J_b at its initial
parameters; on seeded gains and coefficients above the finite-difference
noise floor it must agree per entry to 1e-3 with FP64 central differences of
sum_b J_b (ct_results_objective_f64, h = 1e-6).

The reference runs these scenarios one per thread (scheduler.rs:93,
scenario.rs:797-897); here they share the device.
"""
import hashlib
import sys
import time

import numpy as np
import pytest

from cardiotrust_amd import Algorithm, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

FIELDS = ("gains", "coefs", "delays", "ap_outputs_now", "ap_outputs_last", "system_states",
          "measurements", "d_gains", "d_coefs", "d_coefs_iir", "d_coefs_fir", "loss",
          "loss_mse", "loss_maximum_regularization", "loss_batch", "loss_mse_batch",
          "maximum_regularization_sum", "average_delays", "derivatives_step")
BUILD_WORKERS = 8
H_FD = 1e-6


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    from cardiotrust_amd.gpu import GPU
    g = GPU(0)
    yield g
    g.close()


def digest(r):
    out = r.update_from_gpu(list(FIELDS))
    return {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() for k, v in out.items()}


class Member:
    def __init__(self, gpu, p, cfg, epochs, gradient=None):
        from cardiotrust_amd.gpu import Buffer, EpochKernel, ResultsGPU
        self.r = ResultsGPU(gpu, p, num_batch_entries=epochs)
        self.buf = Buffer(gpu, p.measurements)
        self.k = EpochKernel(gpu, self.r, self.buf, cfg, 3 * p.num_voxels, p.num_sensors,
                             p.num_steps)
        if gradient is not None:
            self.k.set_gradient(*gradient)

    def close(self):
        self.k.close()
        self.buf.close()
        self.r.close()


def simulate_truths(gpu, ps):
    from cardiotrust_amd.gpu import simulate
    for p in ps:
        p.measurements = simulate(gpu, p.truth())


def test_uneven_64_scenario_batch_equals_each_alone(gpu):
    from cardiotrust_amd.gpu import Batch
    t0 = time.perf_counter()
    ps = synth.config3(count=64, workers=BUILD_WORKERS)
    note(f"case: 64 uneven scenarios, {sum(p.num_voxels for p in ps)} voxels, beats "
         f"{sum(p.num_beats for p in ps)}, built in {time.perf_counter() - t0:.0f} s")
    simulate_truths(gpu, ps)
    cfg = Algorithm(learning_rate=1e-3, freeze_delays=False)
    cheapest = int(np.argmin([p.num_voxels * p.num_steps * p.num_beats for p in ps]))
    alone = []
    for i, p in enumerate(ps):
        m = Member(gpu, p, cfg, 2)
        m.k.execute()
        if i == cheapest:
            got = m.r.update_from_gpu(["system_states", "delays", "ap_outputs_now", "gains",
                                       "coefs", "d_gains", "d_coefs", "loss_batch"])
            o = oracle.OracleResults(p, np.float32, num_batch_entries=2)
            o.run_epoch(cfg)
            for k in ("system_states", "delays", "ap_outputs_now"):
                assert np.array_equal(got[k], getattr(o, k)), k
            for k in ("gains", "coefs", "d_gains", "d_coefs", "loss_batch"):
                a, b = got[k].astype(np.float64), getattr(o, k).astype(np.float64)
                assert np.max(np.abs(a - b)) <= 1e-4 * np.max(np.abs(b)), k
            note(f"  scenario {i} ({p.num_voxels} voxels, {p.num_beats} beats, T={p.num_steps}): "
                 f"first epoch matches the oracle")
        m.k.execute()
        alone.append(digest(m.r))
        m.close()
    note(f"  alone: 2 epochs of each scenario done ({time.perf_counter() - t0:.0f} s)")
    members = [Member(gpu, p, cfg, 2) for p in ps]
    b = Batch(gpu, [m.k for m in members], graph=True)
    infos = []
    for _ in range(2):
        b.execute()
        infos.append(b.info())
    b.close()
    note(f"  batch: {infos[-1]['jobs']} jobs in {infos[-1]['rounds']} rounds, voxels per group "
         f"{infos[-1]['voxels_per_group']}, side streams {infos[-1]['side_streams']}, epochs "
         f"{infos[0]['epoch_ms']:.0f} / {infos[1]['epoch_ms']:.0f} ms (graph {infos[1]['graph']})")
    for i, m in enumerate(members):
        got = digest(m.r)
        diff = [f for f in FIELDS if got[f] != alone[i][f]]
        assert not diff, f"scenario {i}: {diff}"
        m.close()
    assert infos[-1]["jobs"] == len(ps)  # one shared propagation per scenario
    assert infos[-1]["side_streams"] == 3
    assert infos[-1]["graph"] == 1


def test_uneven_64_scenario_reverse_mode_batch_equals_each_alone(gpu):
    from cardiotrust_amd.gpu import Batch
    t0 = time.perf_counter()
    ps = synth.config3(count=64, workers=BUILD_WORKERS)
    note(f"case: 64 uneven scenarios in reverse mode (K=16), {sum(p.num_voxels for p in ps)} "
         f"voxels, built in {time.perf_counter() - t0:.0f} s")
    simulate_truths(gpu, ps)
    cfg = Algorithm(learning_rate=1e-3, freeze_delays=False)
    alone = []
    for p in ps:
        m = Member(gpu, p, cfg, 2, gradient=(1, 16))
        m.k.execute()
        m.k.execute()
        alone.append(digest(m.r))
        m.close()
    note(f"  alone: 2 epochs of each scenario done ({time.perf_counter() - t0:.0f} s)")
    members = [Member(gpu, p, cfg, 2, gradient=(1, 16)) for p in ps]
    b = Batch(gpu, [m.k for m in members], graph=True)
    infos = []
    for _ in range(2):
        b.execute()
        infos.append(b.info())
    b.close()
    note(f"  batch: {infos[-1]['jobs']} propagation jobs, {infos[-1]['sweep_jobs']} sweep jobs, "
         f"side streams {infos[-1]['side_streams']}, epochs {infos[0]['epoch_ms']:.0f} / "
         f"{infos[1]['epoch_ms']:.0f} ms (graph {infos[1]['graph']})")
    for i, m in enumerate(members):
        got = digest(m.r)
        diff = [f for f in FIELDS if got[f] != alone[i][f]]
        assert not diff, f"scenario {i}: {diff}"
        m.close()
    assert infos[-1]["sweep_jobs"] == len(ps)  # one backward pass per scenario, one launch
    assert infos[-1]["side_streams"] == 3
    assert infos[-1]["graph"] == 1


def test_shell_adjoint_batch_8x120k_equals_alone_and_fd(gpu):
    from cardiotrust_amd.gpu import Batch
    t0 = time.perf_counter()
    ps = synth.config4(count=8, workers=BUILD_WORKERS)
    note(f"case: 8 shells, voxels {[p.num_voxels for p in ps]}, 4 beats, T=2000, reverse mode "
         f"K=64, built in {time.perf_counter() - t0:.0f} s")
    simulate_truths(gpu, ps)
    grad = (1, 64)
    cfgs = [Algorithm(learning_rate=0.0 if i == 0 else 1e-3, freeze_delays=False)
            for i in range(len(ps))]
    alone = []
    for i, p in enumerate(ps):
        m = Member(gpu, p, cfgs[i], 2, grad)
        for _ in range(2):
            m.k.execute()
        alone.append(digest(m.r))
        m.close()
        note(f"  alone: scenario {i} done ({time.perf_counter() - t0:.0f} s)")
    members = [Member(gpu, p, cfgs[i], 2, grad) for i, p in enumerate(ps)]
    b = Batch(gpu, [m.k for m in members], graph=True)
    infos = []
    for _ in range(2):
        b.execute()
        infos.append(b.info())
    note(f"  batch: epochs {infos[0]['epoch_ms']:.0f} / {infos[1]['epoch_ms']:.0f} ms "
         f"(graph {infos[1]['graph']}), {infos[-1]['jobs']} jobs")
    b.close()
    for i, m in enumerate(members):
        got = digest(m.r)
        diff = [f for f in FIELDS if got[f] != alone[i][f]]
        assert not diff, f"scenario {i}: {diff}"
    assert infos[-1]["graph"] == 1
    # each ~115k-voxel scenario fills the GPU alone: it runs its own resident
    # launch instead of a job of the batched one
    assert infos[-1]["jobs"] == 0

    # member 0 (learning rate 0): its gradient against FP64 central differences
    m, p, cfg = members[0], ps[0], cfgs[0]
    d = m.r.update_from_gpu(["d_gains", "d_coefs", "gains"])
    assert np.array_equal(d["gains"], p.gains)  # lr 0: the parameters did not move
    B = p.num_beats

    def J(param=0, index=0, delta=0.0):
        return sum(m.r.objective_f64(m.buf, beat, cfg, param, index, delta) for beat in range(B))

    J0 = J()
    floor = 1e4 * abs(J0) * np.finfo(np.float64).eps / H_FD
    rng = np.random.default_rng(4)
    for name, param, grad_arr, pool in (
            ("gains", 1, d["d_gains"].ravel(), np.flatnonzero(p.output_state_indices.ravel() >= 0)),
            ("coefs", 2, d["d_coefs"].ravel(), np.flatnonzero(p.geometry.neighbours.ravel() >= 0))):
        resolvable = pool[np.abs(grad_arr[pool]) >= floor]
        picks = rng.choice(resolvable, min(8, resolvable.size), replace=False)
        worst = 0.0
        for idx in picks:
            fd = (J(param, int(idx), H_FD) - J(param, int(idx), -H_FD)) / (2 * H_FD)
            worst = max(worst, abs(grad_arr[idx] - fd) / abs(fd))
        note(f"  {name}: {picks.size} of {resolvable.size} entries above the FD noise floor "
             f"{floor:.1e}: max rel err vs FP64 central differences {worst:.2e} "
             f"({time.perf_counter() - t0:.0f} s)")
        assert picks.size == 8
        assert worst <= 1e-3, name
    for m in members:
        m.close()
