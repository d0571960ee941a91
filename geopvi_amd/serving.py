"""Continuous-batching server of misfit+gradient requests on one GPU.

Requests (a velocity batch of any grid size against observed data kept on
the GPU) arrive from any thread.  The server thread admits queued requests
in arrival order while their device footprint (checkpointed workspace +
outputs) fits the memory budget and the engine has unit slots, advances all
admitted jobs together -- every step launch covers every unit of every job
-- and resolves each request's future as soon as its outputs are final.
Latency is measured from submit() to the moment the result is ready.

This replaces the per-call process pool of the reference's likelihood
(geopvi/forward/fwi2d/posterior.py:131-149): many inversions, each with
many samples, share the device continuously instead of forking a pool per
log_prob call.
"""
from __future__ import annotations

import collections
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from .engine import Engine, EngineJob, best_ckpt_interval, job_bytes
from .forward.fwi2d import solver
from .forward.fwi2d.params import FWIParams


@dataclass
class FWIResult:
    misfit: torch.Tensor      # (nsample,) float64, 0.5*sum((syn-obs)^2)
    grad: torch.Tensor        # (nsample, nz*nx) float32, the solver's imaging gradient (misfit gradient = -grad)
    record_syn: torch.Tensor  # (nsample, ns*nt*nr) float32
    submitted: float
    admitted: float
    finished: float
    extra: dict = field(default_factory=dict)  # "ckpt_interval": the store mode the job ran with (0 = full)

    @property
    def latency(self) -> float:
        return self.finished - self.submitted


def choose_store(full_bytes: int, ckpt_bytes: int, in_use: int, budget: int, policy: str = "auto",
                 full_share: int = 8) -> bool:
    """True: run the job with the full wavefield store; False: checkpointed.

    The full store skips the recompute pass (two passes over the time axis
    instead of three) and gives identical results.  "auto" takes it when the
    job's full-store footprint is at most 1/full_share of the budget and fits
    in what is free now, so a few large jobs cannot crowd out concurrency;
    "full" whenever it fits now; "checkpoint" never.  The caller has checked
    that the checkpointed footprint fits."""
    if policy not in ("auto", "full", "checkpoint"):
        raise ValueError(f"store policy {policy!r}: 'auto', 'full' or 'checkpoint'")
    if policy == "checkpoint":
        return False
    if full_bytes <= ckpt_bytes:  # checkpointing would save nothing
        return True
    fits = in_use + full_bytes <= budget
    if policy == "full":
        return fits
    return fits and full_bytes * full_share <= budget


@dataclass
class _Request:
    params: FWIParams
    vel: torch.Tensor
    obs: torch.Tensor
    run_fwi: int
    ckpt: Optional[int]  # None: chosen at admission by the store policy
    need: int
    ready: torch.cuda.Event
    future: Future
    submitted: float
    admitted: float = 0.0
    job: Optional[EngineJob] = None
    extra: dict = field(default_factory=dict)


class FWIServer:
    """Serve FWI-2D misfit+gradient requests with continuous batching.

    mem_budget: bytes of device memory admitted jobs may hold (default: 90 %
    of the free memory when the server starts, counting blocks torch's
    allocator holds cached; they go back to the device when the server
    closes).  steps_per_tick: time steps
    advanced between admission/completion checks (rounded to whole graph
    segments by the engine).  store_policy: how a request without an explicit
    ckpt_interval keeps its forward wavefield (see choose_store); requests
    are admitted by the footprint of the chosen mode.
    """

    def __init__(self, device: int = 0, mem_budget: Optional[int] = None, steps_per_tick: int = 64,
                 max_units: int = 16384, max_tiles: int = 1 << 21, store_policy: str = "auto"):
        choose_store(0, 0, 0, 1, store_policy)  # validates the name
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        self.stream = torch.cuda.Stream(self.device)
        if mem_budget is None:
            # free device memory plus what torch's allocator holds cached
            free, _ = torch.cuda.mem_get_info(self.device)
            cached = torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
            mem_budget = int((free + cached) * 0.9)
        self.mem_budget = int(mem_budget)
        self.store_policy = store_policy
        self.steps_per_tick = int(steps_per_tick)
        self.engine = Engine(self.stream, max_units=max_units, max_tiles=max_tiles)
        self._queue: collections.deque = collections.deque()
        self._inflight: dict = {}
        self._mem_in_use = 0
        self._units_in_use = 0
        self._cv = threading.Condition()
        self._stop = False
        self._error: Optional[BaseException] = None
        self.completed: list = []
        self._thread = threading.Thread(target=self._run, name="fwi-server", daemon=True)
        self._thread.start()

    # -- client side -----------------------------------------------------------
    def submit(self, params: FWIParams, vel, record_obs, run_fwi: int = 1, ckpt_interval: Optional[int] = None
               ) -> Future:
        """Queue one request; returns a Future of FWIResult.

        vel: (nsample, nz*nx) models (device tensor kept as is, host arrays
        are copied); record_obs: (ns*nt*nr,) observed data on the device."""
        solver.validate(params)
        t_sub = time.perf_counter()
        if not isinstance(vel, torch.Tensor) or not vel.is_cuda:
            vel = torch.as_tensor(np.asarray(vel, np.float32)).to(self.device)
        vel = vel.to(torch.float32).reshape(-1, params.ni).contiguous()
        obs = record_obs if isinstance(record_obs, torch.Tensor) else torch.as_tensor(record_obs)
        obs = obs.to(self.device, torch.float32).reshape(-1).contiguous()
        if obs.numel() != params.record_len:
            raise ValueError(f"record_obs has {obs.numel()} values, expected {params.record_len}")
        # without an explicit interval the store mode is chosen at admission;
        # the checkpointed footprint (the smallest) must fit the budget
        ckpt = None if ckpt_interval is None else int(ckpt_interval)
        need = job_bytes(params, vel.shape[0], best_ckpt_interval(params) if ckpt is None else ckpt)
        if need > self.mem_budget:
            raise MemoryError(f"request needs {need / 1e9:.1f} GB, budget is {self.mem_budget / 1e9:.1f} GB")
        ev = torch.cuda.Event()
        ev.record()  # inputs written on the caller's stream are ready here
        fut: Future = Future()
        req = _Request(params, vel, obs, int(run_fwi), ckpt, need, ev, fut, t_sub)
        with self._cv:
            if self._error is not None:
                raise RuntimeError("server stopped after an error") from self._error
            self._queue.append(req)
            self._cv.notify()
        return fut

    def close(self, wait: bool = True):
        with self._cv:
            self._stop = True
            self._cv.notify()
        if wait:
            self._thread.join()
        self.engine.close()
        if wait:
            torch.cuda.empty_cache()  # job workspaces released by the server thread

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def stats(self) -> dict:
        lat = np.array([r.latency for r in self.completed])
        if lat.size == 0:
            return {"completed": 0}
        return {"completed": int(lat.size), "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean())}

    # -- server thread ---------------------------------------------------------
    def _admit(self):
        while self._queue:
            req = self._queue[0]
            units = req.vel.shape[0] * req.params.ns
            if self._inflight and (self._mem_in_use + req.need > self.mem_budget
                                   or self._units_in_use + units > self.engine.max_units):
                return
            ckpt, need = req.ckpt, req.need
            if ckpt is None:
                full = job_bytes(req.params, req.vel.shape[0], 0)
                if choose_store(full, need, self._mem_in_use, self.mem_budget, self.store_policy):
                    ckpt, need = 0, full
                else:
                    ckpt = best_ckpt_interval(req.params)
            try:
                job = EngineJob(req.params, req.vel, req.obs, ckpt, req.run_fwi)
            except torch.OutOfMemoryError:
                # The budget is counted in bytes, but the caching allocator
                # can be too fragmented to place the workspace: wait for
                # running jobs to release theirs, or, with none running,
                # return cached blocks and try once more.
                if self._inflight:
                    return
                torch.cuda.empty_cache()
                try:
                    job = EngineJob(req.params, req.vel, req.obs, ckpt, req.run_fwi)
                except torch.OutOfMemoryError as exc:
                    self._queue.popleft()
                    req.future.set_exception(MemoryError(f"no room for a {need / 1e9:.1f} GB job: {exc}"))
                    continue
            req.ckpt, req.need = ckpt, need
            self._queue.popleft()
            self.stream.wait_event(req.ready)
            self.engine.submit(job)
            req.job = job
            req.admitted = time.perf_counter()
            self._inflight[job.id] = req
            self._mem_in_use += req.need
            self._units_in_use += units

    def _reap(self):
        done = [jid for jid in self._inflight if self.engine.status(jid) == 1]
        for jid in done:
            req = self._inflight.pop(jid)
            self.engine.release(jid)
            self._mem_in_use -= req.need
            self._units_in_use -= req.vel.shape[0] * req.params.ns
            job = req.job
            res = FWIResult(job.misfit, job.grad, job.syn, req.submitted, req.admitted, time.perf_counter())
            res.extra["ckpt_interval"] = req.ckpt
            job.workspace = None  # give the workspace back to the allocator
            self.completed.append(res)
            req.future.set_result(res)

    def _run(self):
        torch.cuda.set_device(self.device)
        try:
            with torch.cuda.stream(self.stream):
                while True:
                    with self._cv:
                        while not self._stop and not self._queue and not self._inflight:
                            self._cv.wait()
                        if self._stop and not self._inflight and not self._queue:
                            break
                        self._admit()
                    if self._inflight:
                        self.engine.advance(self.steps_per_tick)
                        self._reap()
        except BaseException as exc:  # surface the failure to every waiter
            self._error = exc
            for req in list(self._queue) + list(self._inflight.values()):
                if not req.future.done():
                    req.future.set_exception(exc)
            raise
