#!/usr/bin/env python3
"""Benchmark of the FWI-2D misfit+gradient hot path on one MI355X.

One "step" = one batched misfit+gradient evaluation of the c1 workload:
16 velocity samples (mean-field Gaussian, std 50 m/s around the smoothed
true model) x 20 shots on a 200x400 grid (228x428 padded), 200 receivers,
8 Hz Ricker, dt = 1 ms, 2000 steps, nt_interval = 2, fp32, every
(sample, shot) unit advanced by one launch per time step, replayed as HIP
graph segments.  Inputs (models, observed data) are resident in HBM before
the timed region.

Printed (one JSON line): batched samples/s (``value``), the plain
one-sample-per-launch GPU solver's samples/s on the same workload and the
ratio, the gradient agreement of the two, the roofline of the dominant
kernel (the multi-unit step), and the CPU oracle timed on this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "misfit+grad samples/sec batched vs plain GPU @1e-4 grad tol; HBM GB/s frac; p50/p99"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsample", type=int, default=16)
    ap.add_argument("--ckpt", type=int, default=0, help="checkpoint interval (0 = full wavefield store)")
    ap.add_argument("--plain-samples", type=int, default=1)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-shots", type=int, default=4)
    ap.add_argument("--seed", type=int, default=1)
    return ap.parse_args()


def algorithmic_bytes_per_unit(p) -> int:
    """Bytes a (sample, shot) unit must move at minimum: forward and adjoint
    read p0, p1, v and write u per padded point per step (16 B each pass),
    the wavefield store writes and the imaging condition reads one inner
    field every nt_interval steps and read-modify-writes the gradient
    (16 B per inner point), records/observed/residual 12 B per receiver
    sample."""
    nsave = math.ceil(p.nt / p.nt_interval)
    return p.nt * p.ntp * 32 + nsave * p.ni * 16 + p.nt * p.nr * 12


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(whole (sample, shot) units per GPU, one gradient reduce per sample)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from geopvi_amd import synth, _lib
    from geopvi_amd.forward.fwi2d import solver

    _lib.load()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    p = synth.workload_params("c1")
    rng_seed = args.seed
    vtrue = synth.true_model("c1", seed=rng_seed)
    vmean = synth.smooth(vtrue, 8.0)
    samples = synth.gaussian_samples(vmean, 50.0, args.nsample, seed=rng_seed + 100).astype(np.float32)
    vel = torch.from_numpy(samples).to(dev)
    obs_clean = solver.forward_batch(p, torch.from_numpy(vtrue.astype(np.float32).reshape(1, -1)).to(dev))[0]
    obs_np = synth.add_noise(obs_clean.cpu().numpy(), 0.01, seed=rng_seed + 200)
    obs = torch.from_numpy(obs_np).to(dev)

    wsb = solver.batch_workspace_bytes(p, args.nsample, args.ckpt)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def run_batch():
        return solver.fwi_batch(p, vel, obs, ckpt_interval=args.ckpt, workspace=ws)

    for _ in range(args.warmup):
        run_batch()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for _ in range(args.steps):
        syn_b, grad_b, mis_b = run_batch()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = wall * 1e3 / args.steps
    value = args.nsample * args.steps / wall

    # Roofline of the dominant kernel (multi-unit step): algorithmic bytes of
    # one launch / average launch time.  Launches per batch: nt forward +
    # nt adjoint (+ nt recompute with checkpoints).
    units = args.nsample * p.ns
    launches = p.nt * (2 if args.ckpt <= 0 or args.ckpt >= p.nt else 3)
    bytes_per_launch = units * algorithmic_bytes_per_unit(p) / launches
    avg_launch_s = gpu_ms * 1e-3 / (args.steps * launches)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic = None
    tpath = os.path.join(REPO, "profiles", "traffic_step_multi_c1.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded gradient+anomaly model, Gaussian samples, noisy modelled data)",
        "config": {"workload": "c1: 16 samples x 20 shots, 200x400 grid (228x428 padded), 200 receivers, "
                               "2000 steps, dt=1ms, nt_interval=2, fp32, batched forward+adjoint",
                   "nsample": args.nsample, "shots": p.ns, "nt": p.nt, "grid": [p.nz, p.nx],
                   "ckpt_interval": args.ckpt},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "step_multi_kernel", "bytes_per_launch": int(bytes_per_launch),
                     "avg_launch_us": round(avg_launch_s * 1e6, 3)},
    }

    if not args.no_plain:
        vs0 = vel[0].contiguous()
        solver.fwi_plain(p, vs0, obs)  # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.plain_samples):
            syn_p, grad_p = solver.fwi_plain(p, vel[i % args.nsample].contiguous(), obs)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t0
        plain_sps = args.plain_samples / tp
        gb = grad_b[(args.plain_samples - 1) % args.nsample].double()
        gp = grad_p.double()
        out["plain_samples_per_s"] = round(plain_sps, 4)
        out["speedup_vs_plain"] = round(value / plain_sps, 2)
        out["grad_rel_l2_batched_vs_plain"] = float((gb - gp).norm() / gp.norm())
        mp = 0.5 * float(((syn_p.double() - obs.double()) ** 2).sum())
        mb = float(mis_b[(args.plain_samples - 1) % args.nsample])
        out["misfit_rel_batched_vs_plain"] = abs(mb - mp) / abs(mp)

    if not args.no_cpu:
        import oracle
        ncores = len(os.sched_getaffinity(0))
        oracle.set_threads(ncores)
        shots = max(1, min(args.cpu_shots, p.ns))
        v0 = samples[0].astype(np.float32)
        t0 = time.perf_counter()
        oracle.fwi(p, v0, obs_np, shots=(0, shots))
        tc = time.perf_counter() - t0
        per_sample = tc * p.ns / shots
        out["cpu_baseline"] = {"value": round(1.0 / per_sample, 6), "unit": "samples/s", "cores": ncores,
                               "kind": "port",
                               "sample": f"{shots} of {p.ns} shots of one c1 sample through the C oracle "
                                         f"(OpenMP, reference loop order): {tc:.2f} s, scaled to a sample"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
