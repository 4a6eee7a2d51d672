#!/usr/bin/env python3
"""Epoch throughput of the gfx950 CardioTrust hot path on one MI355X.

Workload: one plain epoch (forward propagation, projection, residuals,
gain and coefficient derivatives, SGD update) of a 64^3 grid masked to an
ellipsoidal shell: N = 120,120 voxels (360,360 states), 26 connections per
voxel with delays 1-40, 108 sensors, T = 2000 steps at 2 kHz, one beat,
learning rate 1e-3, delays not frozen. Inputs are synthetic (seeded); the
measurements are the noise-free GPU simulation of the ground truth.

Prints one JSON line; per-phase timings go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "epochs/s at 120k voxels x 2000 steps (fwd+grad+update), HBM GB/s vs peak"
HBM_PEAK_GBS = 8000.0
MFMA_FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix peak
# Algorithmic bytes of one propagation step per active voxel: 4 B x (234 gains
# + 26 coefs + 26 delays + 2 x 78 filter outputs read/write + 3 states written
# + 3 states read).
PROP_BYTES_PER_VOXEL_STEP = 4 * (234 + 26 + 26 + 2 * 78 + 3 + 3)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(problem, measurements, sample_steps: int, threads_cap: int = 16):
    """Oracle FP32 epoch (single thread) on the first `sample_steps` steps of
    the same problem, scaled to epochs/s of the full T."""
    from oracle import oracle
    from cardiotrust_amd import Algorithm
    from cardiotrust_amd.synth import Problem

    p = Problem(problem.geometry, problem.num_sensors, sample_steps, 1, problem.gains,
                problem.output_state_indices, problem.coefs, problem.delays,
                problem.initial_delays, problem.control_matrix,
                problem.control_function[:sample_steps], problem.measurement_matrix,
                np.ascontiguousarray(measurements[:, :sample_steps]))
    cfg = Algorithm(learning_rate=1e-3, freeze_delays=False)
    r = oracle.OracleResults(p, np.float32)
    t0 = time.perf_counter()
    r.run_epoch(cfg)
    dt = time.perf_counter() - t0
    per_epoch = dt * problem.num_steps / sample_steps
    out = {
        "value": 1.0 / per_epoch,
        "unit": "epochs/s",
        "cores": 1,
        "kind": "port",
        "sample": f"oracle FP32 run_epoch on the first {sample_steps} of {problem.num_steps} steps "
                  f"of the same problem ({dt:.1f} s), scaled by T/{sample_steps}",
        "cpu_model": cpu_model(),
    }
    # throughput variant: one independent scenario epoch per core, concurrently
    # (the reference runs one scenario per thread); the oracle's ctypes calls
    # release the GIL
    import threading
    threads = max(1, min(os.cpu_count() or 1, threads_cap))
    copies = [oracle.OracleResults(p, np.float32) for _ in range(threads)]
    workers = [threading.Thread(target=c.run_epoch, args=(cfg,)) for c in copies]
    t0 = time.perf_counter()
    for w in workers:
        w.start()
    for w in workers:
        w.join()
    dt_all = time.perf_counter() - t0
    out["throughput"] = {
        "value": threads / (dt_all * problem.num_steps / sample_steps),
        "unit": "epochs/s", "cores": threads,
        "sample": f"{threads} concurrent copies of the same sample epoch ({dt_all:.1f} s)",
    }
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_batched(args) -> int:
    """Many scenarios per epoch through ct_batch (one persistent propagation
    launch per round of beats, autotuned voxels per workgroup, hipGraph
    replay), timed beside the same scenarios run one after another."""
    from cardiotrust_amd import Algorithm, synth
    from cardiotrust_amd.gpu import GPU, Batch, Buffer, EpochKernel, ResultsGPU

    t0 = time.perf_counter()
    adjoint = args.workload == "batched-adjoint"
    shared = args.share_propagation == "on" or (args.share_propagation == "auto" and not adjoint)
    if adjoint:
        problems = synth.config4(count=args.scenarios or 8, num_steps=args.num_steps)
        label = (f"{len(problems)} ellipsoidal shells of ~120k voxels with slow sub-regions, "
                 f"4 beats, T={args.num_steps}, reverse-mode gradient (checkpoint every "
                 f"{args.checkpoint} steps), SGD lr 1e-3")
    else:
        problems = synth.config3(count=args.scenarios or 64)
        label = (f"{len(problems)} blob scenarios, voxels log-uniform 1k-40k, 1-8 beats, "
                 f"T uniform 600-2000, reference derivation, SGD lr 1e-3")
    log(f"{len(problems)} synthetic scenarios ({time.perf_counter() - t0:.1f} s)")
    gpu = GPU(0)
    for p in problems:
        tr = ResultsGPU(gpu, p.truth())
        p.measurements = tr.simulate()
        tr.close()
    log(f"ground truths simulated on GPU ({time.perf_counter() - t0:.1f} s)")
    cfg = Algorithm(learning_rate=1e-3, freeze_delays=False)
    epochs = args.warmup + args.steps + 2
    members = []
    for p in problems:
        r = ResultsGPU(gpu, p, num_batch_entries=epochs + 1)
        b = Buffer(gpu, p.measurements)
        k = EpochKernel(gpu, r, b, cfg, 3 * p.num_voxels, p.num_sensors, p.num_steps)
        if adjoint:
            k.set_gradient(1, args.checkpoint)
        k.set_share_propagation({"auto": None, "on": True, "off": False}[args.share_propagation])
        members.append((r, b, k))
    voxel_steps = sum(p.num_voxels * p.num_steps * p.num_beats for p in problems)

    # the same scenarios one after another (one epoch each)
    gpu.synchronize()
    ts = time.perf_counter()
    for r, b, k in members:
        k.execute()
    gpu.synchronize()
    seq_s = time.perf_counter() - ts
    log(f"sequential epoch of all scenarios: {seq_s * 1e3:.1f} ms")

    batch = Batch(gpu, [k for _, _, k in members], voxels_per_group=args.voxels_per_group,
                  graph=not args.no_graph)
    batch.set_profiling(True)
    batch.execute()  # eager epoch (autotune, plan, buffers)
    info0 = batch.info()
    log("autotune (voxels per workgroup -> ms of round 0): " + json.dumps(info0["autotune"]))
    log(f"eager batched epoch: {info0['epoch_ms']:.1f} ms, batched propagation "
        f"{info0['propagate_ms']:.1f} ms in {info0['propagate_launches']} launches")
    for _ in range(max(args.warmup - 1, 0)):
        batch.execute()
    gpu.synchronize()
    ts = time.perf_counter()
    for _ in range(args.steps):
        batch.execute()
    gpu.synchronize()
    elapsed = time.perf_counter() - ts
    info = batch.info()
    ms_per_epoch = 1e3 * elapsed / args.steps
    # roofline of the batched propagation launches (eager epoch, HIP events)
    prop_ms = info0["propagate_ms"]
    resident_vs = 0  # voxel-steps the batched launches propagated in one epoch
    for (r, _, _), p in zip(members, problems):
        if r.set_propagation(0):
            launches = 1 if shared else p.num_beats  # batch_size 0
            resident_vs += p.num_voxels * p.num_steps * launches
    achieved = PROP_BYTES_PER_VOXEL_STEP * resident_vs / (prop_ms * 1e-3) / 1e9 if prop_ms else 0.0
    out = {
        "metric": "batched epochs/s (all scenarios advance one epoch)",
        "value": 1000.0 / ms_per_epoch,
        "unit": "epochs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_epoch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": label, "scenarios": len(problems),
                   "shared_propagation": shared, "share_mode": args.share_propagation,
                   "voxel_steps_per_epoch": voxel_steps, "graph": bool(info["graph"]),
                   "voxels_per_group": info["voxels_per_group"], "rounds": info["rounds"],
                   "jobs": info["jobs"]},
        "scenario_epochs_per_s": len(problems) * 1000.0 / ms_per_epoch,
        "voxel_steps_per_s": voxel_steps * 1000.0 / ms_per_epoch,
        "sequential": {"ms_per_epoch": seq_s * 1e3,
                       "speedup_of_batch": seq_s * 1e3 / ms_per_epoch},
        "roofline": {"bound": "hbm", "kernel": "propagate_batch_kernel", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": None, "propagate_ms_per_epoch": prop_ms,
                     "predicted_ms": info0["predicted_propagate_ms"]},
    }
    batch.close()
    for r, b, k in members:
        k.close()
        b.close()
        r.close()
    print(json.dumps(out), flush=True)
    gpu.close()
    return 0


def bench_copy(args) -> int:
    """Counter calibration: a plain elementwise pass (y = -x, PyTorch's own
    kernel) over a buffer of known size, args.copy_mib MiB read and the same
    written per dispatch. rocprofv3 FETCH_SIZE / WRITE_SIZE of this command
    give the factors that turn the counters of the epoch kernels into bytes."""
    import torch

    n = (args.copy_mib << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.neg(src, out=dst)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        torch.neg(src, out=dst)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    print(json.dumps({"workload": "copy calibration", "mib_read": args.copy_mib,
                      "mib_written": args.copy_mib, "dispatches": args.steps + 1,
                      "ms_per_dispatch": ms,
                      "read_plus_write_gbs": 2 * n * 4 / (ms * 1e-3) / 1e9}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed epochs")
    ap.add_argument("--warmup", type=int, default=2, help="untimed epochs")
    ap.add_argument("--num-steps", type=int, default=2000, help="time steps per beat")
    ap.add_argument("--cpu-sample-steps", type=int, default=50)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--propagation", choices=["auto", "per-step", "resident"], default="auto")
    ap.add_argument("--gradient", choices=["reference", "adjoint"], default="reference",
                    help="reference: the reference's per-step derivation; adjoint: exact "
                         "reverse-mode gradient with checkpoints")
    ap.add_argument("--checkpoint", type=int, default=64, help="adjoint checkpoint interval")
    ap.add_argument("--workload", choices=["plain", "batched-uneven", "batched-adjoint", "copy"],
                    default="plain",
                    help="plain: the 120k-voxel epoch (the headline line); batched-uneven: 64 "
                         "uneven scenarios per batched epoch; batched-adjoint: 8 shells x 4 "
                         "beats with the reverse-mode gradient; copy: HBM counter calibration "
                         "(a plain elementwise pass of --copy-mib MiB)")
    ap.add_argument("--copy-mib", type=int, default=1024, help="copy workload: MiB per dispatch")
    ap.add_argument("--scenarios", type=int, default=0, help="batched workloads: scenario count")
    ap.add_argument("--voxels-per-group", type=int, default=0,
                    help="batched workloads: 0 = autotune")
    ap.add_argument("--no-graph", action="store_true", help="batched workloads: eager epochs")
    ap.add_argument("--share-propagation", choices=["auto", "on", "off"], default="auto",
                    help="batched workloads: one propagation (and, with the reverse-mode "
                         "gradient, one backward pass) per update batch of beats; auto = "
                         "the library default (shared under the reference derivation only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the epoch on a single MI355X; "
              "scenarios would shard one per GPU with no communication.",
              file=sys.stderr)
        return 2
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload != "plain":
        return bench_batched(args)

    from cardiotrust_amd import Algorithm, synth
    from cardiotrust_amd.gpu import GPU, Buffer, EpochKernel, ResultsGPU

    t0 = time.perf_counter()
    problem = synth.config1(num_steps=args.num_steps)
    N, T, S = problem.num_voxels, problem.num_steps, problem.num_sensors
    log(f"synthetic problem: N={N} voxels, T={T}, S={S} ({time.perf_counter() - t0:.1f} s)")
    gpu = GPU(0)
    truth = ResultsGPU(gpu, problem.truth())
    measurements = truth.simulate()
    truth.close()
    problem.measurements = measurements
    log(f"ground truth simulated on GPU ({time.perf_counter() - t0:.1f} s)")

    cfg = Algorithm(learning_rate=1e-3, freeze_delays=False)
    results = ResultsGPU(gpu, problem, num_batch_entries=args.warmup + args.steps + 1)
    mode = {"auto": 0, "per-step": 1, "resident": 2}[args.propagation]
    resident = results.set_propagation(mode)
    log(f"propagation path: {'resident' if resident else 'per-step'}")
    actual = Buffer(gpu, measurements)
    kernel = EpochKernel(gpu, results, actual, cfg, 3 * N, S, T)
    if args.gradient == "adjoint":
        kernel.set_gradient(1, args.checkpoint)
    for _ in range(args.warmup):
        kernel.execute()
    gpu.synchronize()
    log(f"warmup done ({time.perf_counter() - t0:.1f} s)")

    # Timed region: K epochs with per-launch HIP events on the library stream.
    kernel.set_profiling(True)
    prof_sum = {}
    gpu.synchronize()
    start = time.perf_counter()
    for _ in range(args.steps):
        kernel.execute()
        p = kernel.profile()  # synchronises
        for k, v in p.items():
            prof_sum[k] = prof_sum.get(k, 0.0) + v
    gpu.synchronize()
    elapsed = time.perf_counter() - start
    kernel.set_profiling(False)
    ms_per_epoch = 1000.0 * elapsed / args.steps
    prof = {k: v / args.steps for k, v in prof_sum.items()}
    log("per-epoch phases (ms): " + json.dumps({k: round(v, 3) for k, v in prof.items()}))

    launches = max(prof["propagate_launches"], 1)
    launch_ms = prof["propagate_ms"] / launches
    steps_per_launch = prof["propagate_steps"] / launches
    bytes_per_launch = PROP_BYTES_PER_VOXEL_STEP * N * steps_per_launch
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    traffic = None
    tf = os.path.join(ROOT, "profiles", "propagate_resident_cfg1_traffic.json")
    if resident and N == 120120 and os.path.exists(tf):
        # HBM bytes per step from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs
        traffic = json.load(open(tf))["bytes_per_step"] * steps_per_launch
    loss = results.update_from_gpu(["loss_batch"])["loss_batch"]
    log(f"loss_batch per epoch: {loss.tolist()}")

    out = {
        "metric": METRIC,
        "value": 1000.0 / ms_per_epoch,
        "unit": "epochs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_epoch,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plain epoch (fwd+grad+update), 64^3 ellipsoidal shell, N={N} voxels "
                        f"({3 * N} states), 26 neighbours, delays 1-40, S={S} sensors, T={T} "
                        f"steps, 1 beat, SGD lr 1e-3",
            "voxels": N, "steps_per_beat": T, "sensors": S, "beats": 1,
            "propagation": "resident" if resident else "per-step",
            "gradient": args.gradient if args.gradient == "reference"
            else f"adjoint, checkpoint every {args.checkpoint} steps",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "propagate_resident_kernel" if resident else "propagate_step_kernel",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "avg_launch_ms": launch_ms,
            "steps_per_launch": steps_per_launch,
            "us_per_step": 1e3 * launch_ms / steps_per_launch,
            "bytes_per_launch": bytes_per_launch,
        },
    }
    # MFMA utilisation of the two deferred projections (FP32 MFMA, per epoch)
    flop = 2.0 * S * 3 * N * T
    out["mfma"] = {
        "peak_tflops": MFMA_FP32_PEAK_TFLOPS,
        "project": {"kernel": "project_kernel (H·Xᵀ)", "gflop": flop / 1e9,
                    "ms": prof["project_ms"],
                    "tflops": flop / (prof["project_ms"] * 1e-3) / 1e12,
                    "frac": flop / (prof["project_ms"] * 1e-3) / 1e12 / MFMA_FP32_PEAK_TFLOPS},
        "map_residuals": {"kernel": "map_residuals_kernel (Hᵀ·R)", "gflop": flop / 1e9,
                          "ms": prof["map_residuals_ms"],
                          "tflops": flop / (prof["map_residuals_ms"] * 1e-3) / 1e12,
                          "frac": flop / (prof["map_residuals_ms"] * 1e-3) / 1e12
                          / MFMA_FP32_PEAK_TFLOPS},
    }
    kernel.close()
    actual.close()
    results.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(problem, measurements, args.cpu_sample_steps)
    print(json.dumps(out), flush=True)
    gpu.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
