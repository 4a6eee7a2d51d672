"""Writes tests/golden/cfg1_epoch.npz: the oracle's plain epoch at the
120,120-voxel ellipsoidal shell over all T = 2000 steps, as compact vectors
for tests/test_gpu_cfg1_oracle.py.

Run in the build container (no GPU; about 15 minutes on three cores):

    python tests/golden/make_cfg1_fixtures.py

Inputs: synth.config1() (the seeded 64^3 shell: N = 120,120 voxels, delays
1-40, 108 sensors, T = 2000), its ground truth simulated by the FP32 oracle
for the measurements, and one epoch with learning rate 1e-3 and delays not
frozen (the headline workload). The FP32 oracle follows the reference's CPU
epoch operation for operation; the FP64 oracle runs beside it as the
yardstick every compared array also reports against.

Outputs (one .npz, compressed):
  inputs: the measurements (1, T, S) the epoch was given, and SHA-256 digests
    of the initial gains, coefficients and delays, so a test can confirm that
    it rebuilt the same problem;
  states: the SHA-256 of every state row x[t] (T, 32 bytes), and the full last
    row x[T-1];
  64 seeded voxels: their filter outputs after the last step, gain and
    coefficient derivatives, and gains / coefficients / delays after the
    update (FP32, and FP64 where it is a float);
  per-step losses (T), batch losses, every 10th step's measurements, the last
    step's residuals.
"""
from __future__ import annotations

import hashlib
import os
import sys
import threading
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cardiotrust_amd import Algorithm, synth  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = os.path.join(HERE, "cfg1_epoch.npz")
NUM_SAMPLED_VOXELS = 64


def sha256(a: np.ndarray) -> np.ndarray:
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(a).tobytes()).digest(), np.uint8)


def sampled_voxels(N: int) -> np.ndarray:
    u = synth.uniform(1, 123, NUM_SAMPLED_VOXELS)
    return np.unique((u * N).astype(np.int64))


def state_rows(v: np.ndarray) -> np.ndarray:
    return (3 * v[:, None] + np.arange(3)[None, :]).ravel()


def main() -> int:
    t0 = time.perf_counter()
    p = synth.config1()
    print(f"problem: N={p.num_voxels}, T={p.num_steps}, S={p.num_sensors} "
          f"({time.perf_counter() - t0:.0f} s)", flush=True)
    p.measurements = np.ascontiguousarray(oracle.simulate_f32(p.truth()), dtype=np.float32)
    print(f"ground truth simulated by the FP32 oracle ({time.perf_counter() - t0:.0f} s)", flush=True)
    cfg = Algorithm(learning_rate=1e-3, freeze_delays=False)
    runs = {np.float32: oracle.OracleResults(p, np.float32),
            np.float64: oracle.OracleResults(p, np.float64)}
    workers = [threading.Thread(target=o.run_epoch, args=(cfg,)) for o in runs.values()]
    for w in workers:
        w.start()
    for w in workers:
        w.join()
    print(f"FP32 and FP64 epochs done ({time.perf_counter() - t0:.0f} s)", flush=True)
    o32, o64 = runs[np.float32], runs[np.float64]
    v = sampled_voxels(p.num_voxels)
    rows = state_rows(v)
    x = o32.system_states
    out = {
        "measurements_hat": p.measurements,
        "gains_in_sha256": sha256(p.gains), "coefs_in_sha256": sha256(p.coefs),
        "delays_in_sha256": sha256(p.delays),
        "state_row_sha256": np.stack([sha256(x[t]) for t in range(p.num_steps)]),
        "state_last_row": x[-1].copy(),
        "voxels": v,
        "delays": o32.delays[v],
        "ap_outputs_now": o32.ap_outputs_now[rows],
        "residuals": o32.residuals,
    }
    for suffix, o in (("", o32), ("_f64", o64)):
        out.update({
            "d_gains" + suffix: o.d_gains[rows], "d_coefs" + suffix: o.d_coefs[v],
            "gains" + suffix: o.gains[rows], "coefs" + suffix: o.coefs[v],
            "loss" + suffix: o.loss, "loss_mse" + suffix: o.loss_mse,
            "loss_maximum_regularization" + suffix: o.loss_maximum_regularization,
            "loss_batch" + suffix: o.loss_batch, "loss_mse_batch" + suffix: o.loss_mse_batch,
            "loss_maximum_regularization_batch" + suffix: o.loss_maximum_regularization_batch,
            "measurements_every_10" + suffix: o.measurements[0, ::10],
        })
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT) / 1e6:.1f} MB, {time.perf_counter() - t0:.0f} s)",
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
