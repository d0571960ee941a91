"""Launch-shape autotuner of the batched FWI-2D step.

The multi-unit step has three launch knobs (process-wide, set through the C
ABI): the strip-segment height (tile variant: rows per wavefront work item),
the grid (0 = one wavefront per work item) and the time steps per launch
(1 or 2).  Which combination is fastest depends on the grid size, the number
of (sample, shot) units and the device, so this module measures it: every
candidate runs a whole misfit+gradient evaluation of the real geometry --
forward pass with the wavefield store every nt_interval steps, then the
adjoint with the imaging condition on the same steps, the mix of launch
kinds a real evaluation has -- shortened to a few dozen time steps, on
synthetic velocities, timed with HIP events; the fastest one is applied and
remembered in a JSON cache keyed by device and workload shape.  Every candidate computes bit-identical results (they only change
which wavefront computes which points), so tuning never changes outputs.
"""
from __future__ import annotations

import dataclasses
import json
import os
import threading
from typing import Callable, Dict, Iterable, List, Optional, Sequence, Tuple

from . import _lib
from .forward.fwi2d.params import FWIParams

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "geopvi_amd", "autotune.json")

Candidate = Tuple[int, int, int]  # (variant, grid, steps per launch)

_lock = threading.Lock()


def shape_key(params: FWIParams, nsample: int, device_name: str) -> str:
    """Cache key: everything that changes the work-item list and its cost."""
    return (f"{device_name}|nx{params.nx}|nz{params.nz}|pml{params.pml0}|Lc{params.Lc}|ns{params.ns}"
            f"|nr{params.nr}|k{params.nt_interval}|m{nsample}")


def candidates(variants: Iterable[int] = range(6), grids: Sequence[int] = (0, 2048, 3072),
               steps_per_launch: Sequence[int] = (1, 2)) -> List[Candidate]:
    return [(int(v), int(g), int(t)) for v in variants for g in grids for t in steps_per_launch]


def select(timings: Dict[Candidate, float]) -> Candidate:
    """Fastest candidate; ties go to the earliest in iteration order."""
    if not timings:
        raise ValueError("no timings")
    best = None
    for cand, us in timings.items():
        if best is None or us < timings[best]:
            best = cand
    return best


def load_cache(path: str = DEFAULT_CACHE) -> dict:
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def store_cache(key: str, entry: dict, path: str = DEFAULT_CACHE) -> None:
    with _lock:
        data = load_cache(path)
        data[key] = entry
        os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def apply(cand: Candidate, graph_steps: int = 32) -> None:
    variant, grid, tb = cand
    _lib.set_tuning(variant, grid, graph_steps)
    _lib.set_time_blocking(tb)


def measure(params: FWIParams, nsample: int, cands: Sequence[Candidate], steps: int = 64, reps: int = 2,
            device=None, timer: Optional[Callable[[Candidate], float]] = None,
            schedule: str = "evaluation") -> Dict[Candidate, float]:
    """Microseconds per time step (one forward or adjoint step of every
    unit) for each candidate.  schedule "evaluation": forward with store and
    adjoint with imaging over `steps` steps (a real evaluation's launch mix);
    "forward": forward modelling only (no store).

    `timer` replaces the GPU measurement (tests of the selection logic)."""
    if schedule not in ("evaluation", "forward"):
        raise ValueError(f"schedule {schedule!r}: 'evaluation' or 'forward'")
    if timer is not None:
        return {c: float(timer(c)) for c in cands}
    import numpy as np
    import torch

    from .forward.fwi2d import solver
    from . import synth

    dev = torch.device("cuda", 0) if device is None else torch.device(device)
    short = dataclasses.replace(params, nt=max(2 * params.nt_interval, min(params.nt, steps)))
    base = synth.gradient_anomaly_model(params.nz, params.nx, params.dx, params.dz, seed=0)
    vel = synth.gaussian_samples(synth.smooth(base, 4.0), 20.0, nsample, seed=1).astype(np.float32)
    vel_t = torch.from_numpy(vel.reshape(nsample, -1)).to(dev)
    if schedule == "evaluation":
        obs = solver.forward_batch(short, vel_t[:1])[0].contiguous()
        passes = 2

        def run():
            solver.fwi_batch(short, vel_t, obs)
    else:
        passes = 1

        def run():
            solver.forward_batch(short, vel_t)
    saved = (_lib.get_tuning(), _lib.get_time_blocking())
    out: Dict[Candidate, float] = {}
    try:
        for cand in cands:
            apply(cand, saved[0][2])
            run()  # capture the graphs, warm the caches
            torch.cuda.synchronize(dev)
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                run()
            e1.record()
            torch.cuda.synchronize(dev)
            out[cand] = e0.elapsed_time(e1) * 1e3 / (reps * passes * short.nt)
    finally:
        (v, g, s), tb = saved
        _lib.set_tuning(v, g, s)
        _lib.set_time_blocking(tb)
    return out


def autotune(params: FWIParams, nsample: int, cands: Optional[Sequence[Candidate]] = None, steps: int = 64,
             cache_path: Optional[str] = DEFAULT_CACHE, force: bool = False, device=None,
             device_name: Optional[str] = None, timer: Optional[Callable[[Candidate], float]] = None,
             schedule: str = "evaluation") -> dict:
    """Pick and apply the fastest launch shape for (params, nsample).

    Returns {"variant", "grid", "tb", "us_per_step", "cached", "timings"}.
    With a cache path, a previous result for the same device and shape is
    applied without measuring (unless `force`)."""
    if device_name is None:
        if timer is not None:
            device_name = "test"
        else:
            import torch
            device_name = torch.cuda.get_device_name(0 if device is None else device)
    key = shape_key(params, nsample, device_name) + f"|{schedule}"
    if cache_path and not force:
        hit = load_cache(cache_path).get(key)
        if hit and all(k in hit for k in ("variant", "grid", "tb")):
            cand = (int(hit["variant"]), int(hit["grid"]), int(hit["tb"]))
            if timer is None:
                apply(cand)
            return {**hit, "cached": True}
    cands = list(cands) if cands is not None else candidates()
    timings = measure(params, nsample, cands, steps=steps, device=device, timer=timer, schedule=schedule)
    best = select(timings)
    if timer is None:
        apply(best)
    entry = {"variant": best[0], "grid": best[1], "tb": best[2], "us_per_step": round(timings[best], 3),
             "schedule": schedule, "steps": int(steps),
             "timings": {f"{v},{g},{t}": round(us, 3) for (v, g, t), us in timings.items()}}
    if cache_path:
        store_cache(key, entry, cache_path)
    return {**entry, "cached": False}
