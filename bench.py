#!/usr/bin/env python3
"""Benchmark of the FWI-2D misfit+gradient hot path on one MI355X.

One "step" = one batched misfit+gradient evaluation of the c1 workload:
16 velocity samples (mean-field Gaussian, std 50 m/s around the smoothed
true model) x 20 shots on a 200x400 grid (228x428 padded), 200 receivers,
8 Hz Ricker, dt = 1 ms, 2000 steps, nt_interval = 2, fp32, every
(sample, shot) unit advanced by each step launch, replayed as HIP graph
segments.  Inputs (models, observed data) are resident in HBM before the
timed region.  The launch shape is chosen by the autotuner (untimed).

Printed (one JSON line): batched samples/s (``value``), the plain
one-sample-per-launch GPU solver's samples/s on the same workload and the
ratio, the gradient agreement of the two, the roofline of the dominant
kernel (the multi-unit step), a served request stream (continuous
batching; p50/p99 latency), ADVI/PSVI iterations on the device, the
pseudo-spectral path, the CPU oracle timed on this host, and the
CPU-runnable c0 case on CPU (one worker and a per-sample fan-out) and GPU.
"""
from __future__ import annotations

import argparse
import collections
import json
import math
import os

# Before anything starts the HIP runtime: 8 hardware queues for this process
# (HIP's default is 4), the setting a deployment of the server uses.  The
# runtime gives a new stream a least-used queue once all are taken, and with
# 4 the two lane streams of the server measured here were given the same one
# (they then run one after the other: 234 instead of 298 samples/s on the
# serving stream); with 8 they get their own.  Overrides the environment.
os.environ["GPU_MAX_HW_QUEUES"] = "8"
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "misfit+grad samples/sec batched vs plain GPU @1e-4 grad tol; HBM GB/s frac; p50/p99"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--tune", default="", help="variant,grid,steps_per_launch: apply instead of autotuning")
    ap.add_argument("--lanes", type=int, default=0, help="streams per batched call (1..4; 0 = the library's default)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="also run a plain device copy of this many MiB (HBM counter calibration)")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=50.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-low-rate", type=float, default=10.0, help="sub-saturation run rate (0 = skip)")
    ap.add_argument("--serve-low-requests", type=int, default=60)
    ap.add_argument("--serve-c3", type=int, default=6,
                    help="stream this many requests of the serving workload's mix (grids up to 400x800, up to "
                         "64 samples x 40 shots x 3000 steps) at --serve-rate; 0 = skip")
    ap.add_argument("--cpu-shots", type=int, default=5, help="shots of one c1 sample per CPU-baseline run")
    ap.add_argument("--cpu-runs", type=int, default=3, help="timed CPU-baseline runs (median) after a warm-up")
    ap.add_argument("--no-vi", action="store_true", help="skip the end-to-end inversion leg")
    ap.add_argument("--vi-iterations", type=int, default=2)
    ap.add_argument("--no-psm", action="store_true", help="skip the pseudo-spectral (laplace_slover=1) leg")
    ap.add_argument("--no-fwd", action="store_true", help="skip the forward-modelling (forward_2D) leg")
    ap.add_argument("--psm-no-hipfft", action="store_true",
                    help="psm leg without its hipFFT comparison (for counter runs of the library's own kernels)")
    ap.add_argument("--psm-nsample", type=int, default=4)
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


def stream_3r1w_gbs(dev, mib: int = 512, reps: int = 10) -> float:
    """HBM rate of a flat streaming kernel that reads three arrays and
    writes one (torch.addcmul over `mib` MiB each), in GB/s."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.rand(n, device=dev)
    b = torch.rand(n, device=dev)
    c = torch.rand(n, device=dev)
    d = torch.empty(n, device=dev)
    torch.addcmul(a, b, c, out=d)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.addcmul(a, b, c, out=d)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / reps
    del a, b, c, d
    return 4 * n * 4 / sec / 1e9


def serve_stream(args, rate, nreq, seed, dev, mix="bench", compare_batches=False, **server_kw):
    """A Poisson stream of misfit+gradient requests through the server.

    Requests follow one of the serving generator's mixes: "bench" (grids
    100x200 and 200x400, 4-16 samples, 5-10 shots, 500-1000 steps) or "c3"
    (grids 100x200 .. 400x800, 4-64 samples, 5-40 shots, 1000-3000 steps);
    velocity samples around a smoothed model of the c0 generator at that
    size, observed data modelled from the unsmoothed model, both prepared
    before the stream starts.  Latency runs from submit() to the result
    being ready."""
    import torch
    from geopvi_amd import synth
    from geopvi_amd.forward.fwi2d import solver
    from geopvi_amd.serving import FWIServer

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(nreq):
        p, m = synth.serving_request(rng, mix=mix)
        vt = synth.gradient_anomaly_model(p.nz, p.nx, p.dx, p.dz, seed=seed + i)
        vel = synth.gaussian_samples(synth.smooth(vt, 8.0), 50.0, m, seed=seed + 1000 + i).astype(np.float32)
        obs = solver.forward_batch(p, torch.from_numpy(vt.astype(np.float32).reshape(1, -1)).to(dev))[0]
        reqs.append((p, torch.from_numpy(vel).to(dev), obs))
    gaps = rng.exponential(1.0 / rate, size=nreq)
    torch.cuda.synchronize()
    total_bytes = sum(m.shape[0] * p.ns * algorithmic_bytes_per_unit(p) for p, m, _ in reqs)
    total_samples = sum(m.shape[0] for _, m, _ in reqs)
    with FWIServer(dev.index or 0, **server_kw) as srv:
        t0 = time.perf_counter()
        t_next = t0
        futs = []
        for (p, vel, obs), gap in zip(reqs, gaps):
            t_next += gap
            delay = t_next - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futs.append(srv.submit(p, vel, obs))
        res = []
        for i, f in enumerate(futs):
            res.append(f.result())
            if mix != "bench":
                log(f"{mix} request {i + 1}/{nreq} done ({res[-1].latency:.1f} s)")
        t_end = max(r.finished for r in res)
        arena_waits = srv.arena_waits
    lat = np.array([r.latency for r in res])
    wait = np.array([r.admitted - r.submitted for r in res])
    span = t_end - t0
    n_full = sum(1 for r in res if r.extra.get("ckpt_interval") == 0)
    per_lane = collections.Counter(lane for r in res for lane in r.extra.get("lanes", [0]))
    n_split = sum(1 for r in res if r.extra.get("split_for_memory"))
    batches = None
    if compare_batches:
        # The same requests as back-to-back batched calls (full store), for
        # the batch path's rate on exactly this work.
        del res, futs
        for p, vel, obs in reqs:
            solver.fwi_batch(p, vel[:1], obs)
        torch.cuda.synchronize()
        t_b = time.perf_counter()
        for p, vel, obs in reqs:
            solver.fwi_batch(p, vel, obs)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t_b
        batches = {"achieved_GBs": round(total_bytes / tb / 1e9, 1), "seconds": round(tb, 3),
                   "samples_per_s": round(total_samples / tb, 3)}
    policy = server_kw.get("store_policy", "auto")
    return {"requests": nreq, "rate_req_s": rate, "store": f"{policy}: {n_full} full store, {nreq - n_full} checkpointed",
            "split_requests": n_split, "requests_on_each_lane": [per_lane[k] for k in sorted(per_lane)],
            "arena_waits": arena_waits,
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 1),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 1),
            "mean_ms": round(float(lat.mean()) * 1e3, 1),
            "admission_wait_p50_ms": round(float(np.percentile(wait, 50)) * 1e3, 1),
            "makespan_s": round(span, 3),
            "samples_per_s": round(total_samples / span, 3),
            "achieved_GBs": round(total_bytes / span / 1e9, 1),
            **({"same_jobs_as_batches": batches} if batches else {})}


def vi_leg(args, dev):
    """End-to-end variational inversion iterations on the c1 problem
    (200x400 grid, 80 000 parameters, 20 shots, 2000 steps) with a
    Real2Constr bound of [1400, 4800] m/s: mean-field ADVI (diagonal Linear)
    and PSVI (Linear with the 12 bands of a 5x5 correlation template), 32
    samples per iteration, everything on the device; the posterior's
    automatic store policy (the full wavefield store when it fits the free
    device memory, else the interval minimising memory).  One untimed
    warm-up iteration, then `vi_iterations` timed ones per method."""
    import torch
    from geopvi_amd import synth
    from geopvi_amd.engine import best_ckpt_interval
    from geopvi_amd.forward.fwi2d import solver
    from geopvi_amd.forward.fwi2d.posterior import Posterior
    from geopvi_amd.nfvi import FlowsBasedDistribution, Linear, Real2Constr, VariationalInversion, psvi_mask_2D
    from geopvi_amd.prior import Uniform
    import configparser

    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    try:
        p = synth.workload_params("c1")
        vt = synth.true_model("c1", seed=args.seed)
        obs = solver.forward_batch(p, torch.from_numpy(vt.astype(np.float32).reshape(1, -1)).to(dev))[0]
        obs_np = obs.cpu().numpy()
        sigma = 0.01 * float(np.sqrt(np.mean(obs_np.astype(np.float64) ** 2)))
        cfg = configparser.ConfigParser()
        names = [("nx", "nx"), ("nz", "nz"), ("pml0", "pml"), ("Lc", "Lc"), ("laplace_slover", "laplace_slover"),
                 ("ns", "ns"), ("nt", "nt"), ("ds", "ds"), ("ns0", "ns0"), ("depths", "depths"),
                 ("depthr", "depthr"), ("nr", "nr"), ("dr", "dr"), ("nr0", "nr0"), ("nt_interval", "nt_interval"),
                 ("dx", "dx"), ("dz", "dz"), ("dt", "dt"), ("f0", "f0")]
        cfg["FWI"] = {k: (repr(getattr(p, n)) if n in ("dx", "dz", "dt", "f0") else str(getattr(p, n)))
                      for n, k in names}
        dim, lo, up = p.ni, 1400.0, 4800.0
        vs = np.clip(synth.smooth(vt, 8.0).reshape(-1), lo + 50, up - 50)
        u0 = np.log(vs - lo) - np.log(up - vs)
        prior = Uniform(lower=np.full(dim, lo), upper=np.full(dim, up))
        post = Posterior(obs_np, cfg, sigma=sigma, log_prior=prior.log_prob,
                         paramfile=os.path.join("/tmp", f"bench_vi_params_{os.getpid()}.txt"), ckpt_interval="auto")
        out = {"problem": "200x400 (80000 parameters), 20 shots, 2000 steps, bounds [1400, 4800]",
               "nsample": 32}
        for kernel in ("diagonal", "structured"):
            mask = psvi_mask_2D(np.ones((5, 5)), dim, nx=p.nx, nz=p.nz)
            extra = [np.zeros(mask.size)] if kernel == "structured" else []
            flows = [Linear(dim, kernel=kernel, mask=mask, param=np.concatenate([u0, np.full(dim, 0.01)] + extra)),
                     Real2Constr(dim, lower=lo, upper=up)]
            q = FlowsBasedDistribution(flows, generator=torch.Generator(device=dev).manual_seed(args.seed)).to(dev)
            vi = VariationalInversion(q, post.log_prob)
            vi.update(lr=0.002, n_iter=1, nsample=32)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            loss = vi.update(lr=0.002, n_iter=args.vi_iterations, nsample=32)
            torch.cuda.synchronize()
            dt_it = (time.perf_counter() - t0) / args.vi_iterations
            name = "advi" if kernel == "diagonal" else "psvi"
            out[name] = {"s_per_iteration": round(dt_it, 3), "samples_per_s": round(32 / dt_it, 2),
                         "loss": [round(v, 1) for v in loss], "ckpt_interval": post.last_ckpt_interval,
                         "checkpointed_adjoint_interval_if_needed": best_ckpt_interval(p)}
            log(f"vi {name}: {dt_it:.3f} s/iteration")
        return out
    finally:
        torch.set_default_dtype(old)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def c0_baselines(args, ncores, dev):
    """The CPU-runnable case (c0: 120x200, 4 shots, 1000 steps, one sample):
    the C oracle with all host cores (median of 3 after a warm-up); the
    reference's per-sample fan-out (posterior.py:143-144 runs one worker per
    sample, each with its own OpenMP threads) emulated as P concurrent
    workers x T threads, P*T = cores (ctypes releases the GIL, and the
    OpenMP thread count is per calling thread); and the plain and batched
    GPU solvers on the same case."""
    import threading
    import torch
    import oracle
    from geopvi_amd import synth
    from geopvi_amd.forward.fwi2d import solver

    p0 = synth.workload_params("c0")
    vt = synth.true_model("c0", seed=args.seed)
    vs = synth.smooth(vt, 8.0).astype(np.float32)
    oracle.set_threads(ncores)
    obs = oracle.forward(p0, vt)
    oracle.fwi(p0, vs, obs)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.fwi(p0, vs, obs)
        ts.append(time.perf_counter() - t0)
    med = float(np.median(ts))
    nproc = 4 if ncores >= 8 else 2
    nthr = max(1, ncores // nproc)
    batch = synth.gaussian_samples(vs, 50.0, nproc, seed=args.seed + 500).astype(np.float32)

    def worker(i):
        oracle.set_threads(nthr)
        oracle.fwi(p0, batch[i], obs)

    t0 = time.perf_counter()
    th = [threading.Thread(target=worker, args=(i,)) for i in range(nproc)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    pool_s = time.perf_counter() - t0
    oracle.set_threads(ncores)
    log(f"c0 cpu: {med:.3f} s/sample on {ncores} threads; fan-out {nproc}x{nthr}: {pool_s:.3f} s")

    vd = torch.from_numpy(vs.reshape(-1)).to(dev)
    od = torch.from_numpy(obs).to(dev)
    solver.fwi_plain(p0, vd, od)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        solver.fwi_plain(p0, vd, od)
    torch.cuda.synchronize()
    plain = (time.perf_counter() - t0) / 3
    vb = torch.from_numpy(synth.gaussian_samples(vs, 50.0, 16, seed=args.seed + 600).astype(np.float32)).to(dev)
    solver.fwi_batch(p0, vb, od)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.fwi_batch(p0, vb, od)
    torch.cuda.synchronize()
    batched = time.perf_counter() - t0
    return {"workload": "c0: 120x200 grid, 4 shots, 100 receivers, 1000 steps",
            "cpu_model": _cpu_model(), "cores": ncores,
            "cpu_oracle_s_per_sample_median3": round(med, 4),
            "cpu_fanout": {"workers": nproc, "threads_each": nthr, "samples": nproc,
                           "samples_per_s": round(nproc / pool_s, 3)},
            "gpu_plain_samples_per_s": round(1.0 / plain, 2),
            "gpu_batched_16_samples_per_s": round(16 / batched, 2)}


def _recorded_traffic():
    """HBM bytes per launch of the c1 step kernel from the profiler run on
    record (profiles/traffic_step_multi_c1.json), with its launch shape."""
    path = os.path.join(REPO, "profiles", "traffic_step_multi_c1.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def forward_leg(dev, p, vel):
    """Forward modelling only (forward_2D, fdtd2D_modelling.c:356-675): the
    c1 records of every sample, batched (no wavefield store), against one
    sample through the shots-in-turn solver.  One untimed warm-up each."""
    import torch
    from geopvi_amd.forward.fwi2d import solver

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    n = vel.shape[0]
    t_batch = timed(lambda: solver.forward_batch(p, vel))
    t_plain = timed(lambda: solver.forward_plain(p, vel[0].contiguous()))
    # per unit and step: read p0, p1 and v, write p2 over the padded grid,
    # plus the receiver row
    ntp = (p.nx + 2 * (p.pml0 + p.Lc + 1)) * (p.nz + 2 * (p.pml0 + p.Lc + 1))
    bytes_per_step = n * p.ns * (16 * ntp + 4 * p.nr)
    out = {"nsample": n, "shots": p.ns, "nt": p.nt, "batched_samples_per_s": round(n / t_batch, 3),
           "plain_samples_per_s": round(1 / t_plain, 4), "speedup_vs_plain": round(t_plain * n / t_batch, 1),
           "us_per_time_step_batched": round(t_batch / p.nt * 1e6, 1),
           "algorithmic_GBs": round(bytes_per_step * p.nt / t_batch / 1e9, 1)}
    log(f"forward: batched {n / t_batch:.2f} samples/s, plain {1 / t_plain:.3f} samples/s")
    return out


def _radices(n):
    """The hand-written transforms' passes for length n (radix 4 first, then
    2, 3, 5, 7, 11, 13), or None when n has another prime factor."""
    out, m = [], n
    while m % 4 == 0:
        out.append(4)
        m //= 4
    for r in (2, 3, 5, 7, 11, 13):
        while m % r == 0:
            out.append(r)
            m //= r
    return out if m == 1 else None


def psm_leg(args, dev, p_fd, samples, obs):
    """The pseudo-spectral Laplacian (laplace_slover = 1) on the c1 geometry:
    `psm_nsample` samples x 20 shots batched (double-precision transforms of
    every unit), one sample alone through the plain entry point, and the
    batch again with hipFFT's transforms (GPVI_PSM_HIPFFT=1) for comparison.
    One untimed warm-up evaluation each."""
    import torch
    from geopvi_amd.forward.fwi2d import solver

    p = p_fd.replace(laplace_slover=1)
    n = max(1, min(args.psm_nsample, len(samples)))
    vel = torch.from_numpy(np.ascontiguousarray(samples[:n], np.float32)).to(dev)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    t_batch = timed(lambda: solver.fwi_batch(p, vel, obs))
    t_plain = timed(lambda: solver.fwi_plain(p, vel[0], obs))
    t_lib = None
    if not args.psm_no_hipfft:
        os.environ["GPVI_PSM_HIPFFT"] = "1"
        try:
            t_lib = timed(lambda: solver.fwi_batch(p, vel, obs))
        finally:
            del os.environ["GPVI_PSM_HIPFFT"]
    ntx, ntz = p.nx + 2 * p.pml0, p.nz + 2 * p.pml0
    rx, rz = _radices(ntx), _radices(ntz)
    own = rx is not None and rz is not None
    out = {"grid_padded": [ntz, ntx], "nsample": n, "shots": p.ns, "nt": p.nt,
           "batched_samples_per_s": round(n / t_batch, 4), "plain_samples_per_s": round(1 / t_plain, 4),
           "speedup_vs_plain": round(t_plain * n / t_batch, 2),
           "us_per_time_step_batched": round(t_batch / (2 * p.nt) * 1e6, 1),
           "transforms": (f"hand-written mixed-radix in LDS, double (rows {'x'.join(map(str, rx))}, columns "
                          f"{'x'.join(map(str, rz))}; two kernels per time step within a pass)") if own else
                         "hipFFT D2Z + Z2D 2-D in place, double, batch = samples x shots, one pair per time step",
           "hipfft_batched_samples_per_s": None if t_lib is None else round(n / t_lib, 4),
           "hipfft_us_per_time_step_batched": None if t_lib is None else round(t_lib / (2 * p.nt) * 1e6, 1)}
    lib_s = "not run" if t_lib is None else f"{n / t_lib:.3f}"
    log(f"psm: batched {n / t_batch:.3f} samples/s (hipFFT {lib_s}), plain {1 / t_plain:.3f} samples/s")
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the solver runs on one GPU; multi-GPU sharding is design-only "
                                   "(whole (sample, shot) units per GPU, one gradient reduce per sample)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    from geopvi_amd import _lib, autotune, synth
    from geopvi_amd.forward.fwi2d import solver

    _lib.load()
    if args.lanes:
        _lib.set_lanes(args.lanes)
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
    del obs_clean

    log("inputs ready")
    if args.copy_calibration_mb > 0:
        # A kernel of known traffic (reads and writes N MiB once) for
        # scaling the profiler's FETCH_SIZE / WRITE_SIZE counters.
        n = args.copy_calibration_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    tune = None
    if args.tune:
        autotune.apply(tuple(int(x) for x in args.tune.split(",")))
    elif not args.no_autotune:
        tune = autotune.autotune(p, args.nsample, cache_path=None)
        # Shapes within 3 % of the fastest are the same speed to the
        # autotuner's resolution (64 forward steps); prefer the one whose HBM
        # counters are on record, so the roofline's traffic figure applies.
        rec = _recorded_traffic()
        if rec:
            key = f"{rec['variant']},{rec['grid']},{rec['tb']}"
            t_rec = tune["timings"].get(key)
            if t_rec is not None and t_rec <= 1.03 * tune["us_per_step"]:
                autotune.apply((rec["variant"], rec["grid"], rec["tb"]))
                tune = {**tune, "applied": key, "note": "recorded-traffic shape within 3 % of the fastest"}
    variant, grid, graph_steps = _lib.get_tuning()
    tb = _lib.get_time_blocking()
    # Streams the batch's units are split over (gpvi_fwi2d_set_lanes): every
    # macro step is then one concurrent dispatch per lane.
    lanes = min(_lib.get_lanes(), args.nsample) if args.nsample >= 2 else min(_lib.get_lanes(), 2)
    log(f"launch shape: variant {variant} grid {grid} steps/launch {tb} lanes {lanes}")

    wsb = solver.batch_workspace_bytes(p, args.nsample, args.ckpt)
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()

    def run_batch():
        return solver.fwi_batch(p, vel, obs, ckpt_interval=args.ckpt, workspace=ws)

    for _ in range(args.warmup):
        run_batch()
    torch.cuda.synchronize()
    # Events on torch's current stream: the stream every kernel of the
    # batched evaluation (graph segments included) is enqueued on.
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
    log(f"timed {args.steps} batched evaluations: {wall:.3f} s")
    ms_per_step = wall * 1e3 / args.steps
    value = args.nsample * args.steps / wall

    # Roofline of the dominant kernel (multi-unit step).  A launch here is
    # one macro step of every unit: `lanes` concurrent dispatches, each over
    # its share of the units, on the lanes' streams.  Algorithmic bytes of
    # one launch / its average time, the timed region's time (events on the
    # caller's stream, which joins the lanes' streams at the end of every
    # evaluation) over the launch count.  The step launches are all but a
    # few small set-up and reduction kernels of the region, so this bounds
    # the step's average time from above; the profile under profiles/ gives
    # the dispatches' union per launch from the kernel trace.
    sch = solver.schedule(p, args.ckpt, tb)
    units = args.nsample * p.ns
    launches = sch["launches"]
    bytes_per_launch = units * algorithmic_bytes_per_unit(p) / launches
    avg_launch_s = gpu_ms * 1e-3 / (args.steps * launches)
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic = None
    tj = _recorded_traffic()
    profile = None
    if (tj and tj.get("tb") == tb and tj.get("variant") == variant and tj.get("ckpt") == args.ckpt
            and tj.get("lanes", 1) == lanes):
        traffic = tj.get("hbm_bytes_per_launch")
        # the profile on record: the union of the lanes' dispatch intervals per
        # launch (comparable with avg_launch_us) and one dispatch's average
        profile = {"union_us_per_launch": tj.get("union_us_per_launch"),
                   "avg_dispatch_us": tj.get("avg_dispatch_us"),
                   "file": "profiles/step_multi_kernel_c1_16x20.md"}

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
                   "ckpt_interval": args.ckpt, "tile_variant": variant, "launch_grid": grid,
                   "steps_per_launch": tb, "graph_steps": graph_steps, "lanes": lanes},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "step_multi_kernel", "bytes_per_launch": int(bytes_per_launch),
                     "launches_per_step": launches, "avg_launch_us": round(avg_launch_s * 1e6, 3),
                     "dispatches_per_launch": lanes, **({"profile_on_record": profile} if profile else {})},
    }
    if tune is not None:
        out["autotune"] = {"schedule": "forward with store + adjoint with imaging, 64 steps",
                           "chosen": tune.get("applied", f"{tune['variant']},{tune['grid']},{tune['tb']}"),
                           "us_per_time_step": tune["us_per_step"], "candidates": len(tune["timings"]),
                           "us_per_launch_full_evaluation": round(avg_launch_s * 1e6, 3)}
    del ws
    # What a plain streaming kernel with the step's read:write mix reaches on
    # this GPU (torch's addcmul: 3 reads, 1 write), for scale.
    stream = stream_3r1w_gbs(dev)
    out["roofline"]["stream_3r1w_GBs"] = round(stream, 1)
    out["roofline"]["frac_of_stream"] = round(achieved / stream, 4)

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
        log(f"plain solver: {tp:.3f} s for {args.plain_samples} sample(s)")
        gb = grad_b[(args.plain_samples - 1) % args.nsample].double()
        gp = grad_p.double()
        out["plain_samples_per_s"] = round(plain_sps, 4)
        out["speedup_vs_plain"] = round(value / plain_sps, 2)
        out["grad_rel_l2_batched_vs_plain"] = float((gb - gp).norm() / gp.norm())
        mp = 0.5 * float(((syn_p.double() - obs.double()) ** 2).sum())
        mb = float(mis_b[(args.plain_samples - 1) % args.nsample])
        out["misfit_rel_batched_vs_plain"] = abs(mb - mp) / abs(mp)
    del syn_b, grad_b
    torch.cuda.empty_cache()

    if not args.no_serve:
        serve = {"mix": "bench (grids 100x200/200x400, 4-16 samples, 5-10 shots, 500-1000 steps)"}
        serve["saturated"] = serve_stream(args, args.serve_rate, args.serve_requests, rng_seed + 300, dev,
                                          compare_batches=True)
        log(f"served stream at {args.serve_rate}/s done")
        if args.serve_low_rate > 0:
            serve["sub_saturation"] = serve_stream(args, args.serve_low_rate, args.serve_low_requests,
                                                   rng_seed + 400, dev)
        out["serving"] = serve
    if args.serve_c3 > 0:
        out["serving_c3"] = {"mix": "c3 (grids 100x200..400x800, 4-64 samples, 5-40 shots, 1000-3000 steps)",
                             **serve_stream(args, args.serve_rate, args.serve_c3, rng_seed + 700, dev, mix="c3")}
        log(f"c3 stream of {args.serve_c3} requests done")
        torch.cuda.empty_cache()

    if not args.no_vi:
        out["vi"] = vi_leg(args, dev)
        torch.cuda.empty_cache()

    if not args.no_psm:
        out["psm"] = psm_leg(args, dev, p, samples, obs)
        torch.cuda.empty_cache()

    if not args.no_fwd:
        out["forward_modelling"] = forward_leg(dev, p, vel)
        torch.cuda.empty_cache()

    if not args.no_cpu:
        import oracle
        ncores = len(os.sched_getaffinity(0))
        oracle.set_threads(ncores)
        shots = max(1, min(args.cpu_shots, p.ns))
        v0 = samples[0].astype(np.float32)
        oracle.fwi(p, v0, obs_np, shots=(0, 1))   # warm-up
        runs = []
        for _ in range(max(1, args.cpu_runs)):
            t0 = time.perf_counter()
            oracle.fwi(p, v0, obs_np, shots=(0, shots))
            runs.append(time.perf_counter() - t0)
        tc = float(np.median(runs))
        log(f"cpu oracle: {shots} shots, median {tc:.2f} s of {len(runs)} runs")
        per_sample = tc * p.ns / shots
        out["cpu_baseline"] = {"value": round(1.0 / per_sample, 6), "unit": "samples/s", "cores": ncores,
                               "kind": "port", "cpu_model": _cpu_model(),
                               "runs_s": [round(t, 3) for t in runs],
                               "sample": f"{shots} of {p.ns} shots of one c1 sample through the C oracle "
                                         f"(OpenMP on {ncores} threads, reference loop order, full per-call "
                                         f"set-up); median of {len(runs)} runs after a one-shot warm-up: "
                                         f"{tc:.2f} s, scaled to the sample's {p.ns} shots"}
        out["c0_baselines"] = c0_baselines(args, ncores, dev)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
