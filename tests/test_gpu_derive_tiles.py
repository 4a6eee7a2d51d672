"""The derivation over voxel tiles (ct_results_set_derivation 2, the default
when the tiles fit) against a lane per connection (1), bit for bit; and the
same for the reverse-mode gradient pass at checkpoints every 16 steps.

Both compute each connection's filter outputs, FIR / IIR sensitivities and
derivative sums with the same operations in the same order; they differ only
in where the delayed neighbour states come from (an LDS ring filled once per
tile, or memory). So every result array of an epoch must be the same bits.
The cases cover a partial last chunk of steps (T not a multiple of 16),
delays above the ring's 48 steps (those connections read memory), missing
neighbours, both coefficient derivatives, frozen gains or delays, the
maximum and difference regularisers, two beats with two updates per epoch,
and the 120,120-voxel shell at T = 2000. The oracle parity tests run the
default path, so they check the tiled kernel against the oracle as well.

The gradient pass over tiles walks each tile's chunks backwards with the ring
filled in that order; its cases add fewer than four chunks, delays above its
47-step limit, and two beats under one backward pass (shared propagation).
"""
import hashlib
import sys

import numpy as np
import pytest

from cardiotrust_amd import Algorithm, APDerivative, synth
from oracle import oracle

pytestmark = pytest.mark.gpu

FIELDS = ("gains", "coefs", "delays", "ap_outputs_now", "ap_outputs_last", "system_states",
          "measurements", "d_gains", "d_coefs", "d_coefs_iir", "d_coefs_fir", "loss",
          "loss_mse", "loss_maximum_regularization", "loss_batch", "loss_mse_batch",
          "maximum_regularization_sum", "average_delays")


def note(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    from cardiotrust_amd.gpu import GPU
    g = GPU(0)
    yield g
    g.close()


def run(gpu, p, cfg, mode, epochs=1, nbatch=1, gradient=None):
    from cardiotrust_amd.gpu import Buffer, EpochKernel, ResultsGPU
    r = ResultsGPU(gpu, p, num_batch_entries=nbatch)
    tiled = r.set_derivation(mode)
    assert tiled == (mode != 1)
    buf = Buffer(gpu, p.measurements)
    k = EpochKernel(gpu, r, buf, cfg, 3 * p.num_voxels, p.num_sensors, p.num_steps)
    if gradient is not None:
        k.set_gradient(*gradient)
    for _ in range(epochs):
        k.execute()
    out = r.update_from_gpu(list(FIELDS))
    k.close()
    buf.close()
    r.close()
    return {f: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() for f, v in out.items()}


def blob_problem(seed, shape, T, max_delay, fill=0.6, beats=1):
    mask = synth.uniform(seed, 99, int(np.prod(shape))).reshape(shape) < fill
    mask.flat[0] = True
    g = synth.grid_geometry(mask, 2.5)
    offs = np.stack([np.array([3.0 * b, 0.0, 0.0]) for b in range(beats)])
    return synth.make_problem(g, num_steps=T, max_delay=max_delay, seed=seed, num_beats=beats,
                              gain_sigma=0.1, beat_offsets_mm=offs, simulate=oracle.simulate_f32)


CASES = [
    dict(name="cube 10^3, T=1000, delays 1-20", make=lambda: synth.config0(simulate=oracle.simulate_f32),
         cfg=dict(learning_rate=1e-3, freeze_delays=False)),
    dict(name="blob 9x8x7, T=203, delays 1-60", make=lambda: blob_problem(11, (9, 8, 7), 203, 60),
         cfg=dict(learning_rate=0.3, freeze_delays=False, maximum_regularization_threshold=0.05)),
    dict(name="blob 9x8x7, T=203, delays 1-60, Simple + difference regulariser",
         make=lambda: blob_problem(11, (9, 8, 7), 203, 60),
         cfg=dict(learning_rate=0.3, freeze_delays=False, ap_derivative=APDerivative.Simple,
                  difference_regularization_strength=1e-3)),
    dict(name="blob 8x7x6 fill 0.45, T=77, delays 1-30, frozen gains",
         make=lambda: blob_problem(4, (8, 7, 6), 77, 30, fill=0.45),
         cfg=dict(learning_rate=0.3, freeze_delays=False, freeze_gains=True)),
    dict(name="blob 8x7x6 fill 0.45, T=77, delays 1-30, frozen delays",
         make=lambda: blob_problem(4, (8, 7, 6), 77, 30, fill=0.45),
         cfg=dict(learning_rate=0.3, freeze_delays=True)),
    dict(name="blob 8x7x6, 2 beats, batch 1, T=120, delays 1-48",
         make=lambda: blob_problem(6, (8, 7, 6), 120, 48, beats=2),
         cfg=dict(learning_rate=0.3, freeze_delays=False, batch_size=1), nbatch=2),
]


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_tiled_derivation_equals_lane_per_connection(gpu, case):
    p = case["make"]()
    note(f"case: {case['name']}: N={p.num_voxels}, two epochs, lane per connection vs tiles")
    cfg = Algorithm(**case["cfg"])
    nb = case.get("nbatch", 1)
    lanes = run(gpu, p, cfg, 1, epochs=2, nbatch=2 * nb)
    tiles = run(gpu, p, cfg, 2, epochs=2, nbatch=2 * nb)
    auto = run(gpu, p, cfg, 0, epochs=2, nbatch=2 * nb)
    diff = [f for f in FIELDS if lanes[f] != tiles[f]]
    assert not diff, diff
    assert auto == tiles


def test_tiled_derivation_at_the_120k_shell(gpu):
    p = synth.config1()
    p.measurements = np.zeros((1, p.num_steps, p.num_sensors), np.float32)
    note(f"case: 120k-voxel shell N={p.num_voxels}, T={p.num_steps}, zero measurements, one epoch, "
         f"lane per connection vs tiles")
    cfg = Algorithm(learning_rate=1e-3, freeze_delays=False)
    lanes = run(gpu, p, cfg, 1)
    tiles = run(gpu, p, cfg, 2)
    diff = [f for f in FIELDS if lanes[f] != tiles[f]]
    assert not diff, diff


ADJOINT_CASES = [
    dict(name="cube 10^3, T=1000, delays 1-20", make=lambda: synth.config0(simulate=oracle.simulate_f32)),
    dict(name="blob 9x8x7, T=203, delays 1-60", make=lambda: blob_problem(11, (9, 8, 7), 203, 60)),
    dict(name="blob 8x7x6 fill 0.45, T=50, delays 1-30", make=lambda: blob_problem(4, (8, 7, 6), 50, 30,
                                                                              fill=0.45)),
    dict(name="blob 8x7x6, 2 beats shared, T=120, delays 1-47",
         make=lambda: blob_problem(6, (8, 7, 6), 120, 47, beats=2)),
]


@pytest.mark.parametrize("case", ADJOINT_CASES, ids=[c["name"] for c in ADJOINT_CASES])
def test_tiled_gradient_pass_equals_lane_per_connection(gpu, case):
    p = case["make"]()
    note(f"case: {case['name']}: N={p.num_voxels}, reverse mode K=16, two epochs, lane per "
         f"connection vs tiles")
    cfg = Algorithm(learning_rate=0.3, freeze_delays=False)
    lanes = run(gpu, p, cfg, 1, epochs=2, nbatch=2, gradient=(1, 16))
    tiles = run(gpu, p, cfg, 2, epochs=2, nbatch=2, gradient=(1, 16))
    diff = [f for f in FIELDS if lanes[f] != tiles[f]]
    assert not diff, diff


def test_tiled_gradient_pass_at_the_120k_shell(gpu):
    p = synth.config1()
    p.measurements = np.zeros((1, p.num_steps, p.num_sensors), np.float32)
    note(f"case: 120k-voxel shell N={p.num_voxels}, T={p.num_steps}, reverse mode K=16, one epoch, "
         f"lane per connection vs tiles")
    cfg = Algorithm(learning_rate=1e-3, freeze_delays=False)
    lanes = run(gpu, p, cfg, 1, gradient=(1, 16))
    tiles = run(gpu, p, cfg, 2, gradient=(1, 16))
    diff = [f for f in FIELDS if lanes[f] != tiles[f]]
    assert not diff, diff
