"""A batch at the uneven workload's real sizes against each scenario alone.

The first 12 scenarios of the seeded 64-scenario uneven workload (voxels
log-uniform in 1k-40k, 1-8 beats, 600-2000 steps, delays 1-20) run two epochs
as one batch (graph replay, autotuned slab size) and each on its own epoch
kernel, with one propagation per beat (share mode off) so that the batch
carries one job per beat. Every result array must be the same bits: compared
by SHA-256 so the host holds one member's arrays at a time.
"""
import hashlib
import sys

import numpy as np
import pytest

from cardiotrust_amd import Algorithm, synth

pytestmark = pytest.mark.gpu

FIELDS = ("gains", "coefs", "delays", "ap_outputs_now", "ap_outputs_last", "system_states",
          "measurements", "d_gains", "d_coefs", "d_coefs_iir", "d_coefs_fir", "loss",
          "loss_batch", "loss_mse_batch", "maximum_regularization_sum", "average_delays")


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


def member(gpu, p, cfg, epochs):
    from cardiotrust_amd.gpu import Buffer, EpochKernel, ResultsGPU
    r = ResultsGPU(gpu, p, num_batch_entries=epochs)
    buf = Buffer(gpu, p.measurements)
    k = EpochKernel(gpu, r, buf, cfg, 3 * p.num_voxels, p.num_sensors, p.num_steps)
    k.set_share_propagation(False)  # one propagation job per beat: 61 jobs in all
    return r, buf, k


def test_batch_of_workload_size_scenarios_equals_each_alone(gpu):
    from cardiotrust_amd.gpu import Batch, simulate
    ps = synth.config3(count=12)
    note("case: 12 uneven scenarios, voxels " + ", ".join(str(p.num_voxels) for p in ps) +
         f"; beats {[p.num_beats for p in ps]}; 2 epochs, batch (graph) vs alone")
    for p in ps:
        p.measurements = simulate(gpu, p.truth())
    cfg = Algorithm(learning_rate=1e-3, freeze_delays=False)
    alone = []
    for p in ps:
        r, buf, k = member(gpu, p, cfg, 2)
        for _ in range(2):
            k.execute()
        alone.append(digest(r))
        k.close()
        buf.close()
        r.close()
    members = [member(gpu, p, cfg, 2) for p in ps]
    b = Batch(gpu, [k for _, _, k in members], graph=True)
    for _ in range(2):
        b.execute()
    info = b.info()
    note(f"  batch: {info['jobs']} jobs in {info['rounds']} rounds, voxels per group "
         f"{info['voxels_per_group']}, graph {info['graph']}")
    b.close()
    for i, (r, buf, k) in enumerate(members):
        got = digest(r)
        diff = [f for f in FIELDS if got[f] != alone[i][f]]
        assert not diff, f"scenario {i}: {diff}"
        k.close()
        buf.close()
        r.close()
    assert info["jobs"] == sum(p.num_beats for p in ps)
