"""Continuous-batching server of misfit+gradient requests on one GPU.

Requests (a velocity batch of any grid size against observed data kept on
the GPU) arrive from any thread.  The server thread admits queued requests
in arrival order while their workspace (full wavefield store or
checkpoints, see choose_store) can be placed in the server's arena and the
engine has unit slots, advances all admitted jobs together -- every step
launch covers every unit of every job -- and resolves each request's future
as soon as its outputs are final.  Latency is measured from submit() to the
moment the result is ready.

The arena is one device allocation of the whole memory budget, made when
the server starts; job workspaces are carved out of it by a host-side
first-fit allocator, so admission is decided by the arena's free blocks and
no device allocation or release (and none of the device-wide
synchronisation that comes with them) happens while the server runs.  The
small per-job outputs are ordinary tensors handed to the caller.

This replaces the per-call process pool of the reference's likelihood
(geopvi/forward/fwi2d/posterior.py:131-149): many inversions, each with
many samples, share the device continuously instead of forking a pool per
log_prob call.
"""
from __future__ import annotations

import bisect
import collections
import json
import math
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib
from .engine import Engine, EngineJob, best_ckpt_interval, job_workspace_bytes
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


class Arena:
    """First-fit allocator of 256-byte-aligned blocks over one device buffer
    (host-side bookkeeping; the buffer is allocated once)."""

    ALIGN = 256

    def __init__(self, nbytes: int, device: torch.device):
        self.size = int(nbytes) // self.ALIGN * self.ALIGN
        self.buf = torch.empty(max(self.size, self.ALIGN), dtype=torch.uint8, device=device)
        self._free = [(0, self.size)]   # sorted, coalesced (offset, size)

    @classmethod
    def _round(cls, n: int) -> int:
        return (int(n) + cls.ALIGN - 1) // cls.ALIGN * cls.ALIGN

    def largest(self) -> int:
        return max((sz for _, sz in self._free), default=0)

    def free_bytes(self) -> int:
        return sum(sz for _, sz in self._free)

    def alloc(self, nbytes: int) -> Optional[int]:
        n = self._round(nbytes)
        for i, (off, sz) in enumerate(self._free):
            if sz >= n:
                if sz == n:
                    del self._free[i]
                else:
                    self._free[i] = (off + n, sz - n)
                return off
        return None

    def release(self, off: int, nbytes: int) -> None:
        n = self._round(nbytes)
        i = bisect.bisect_left(self._free, (off, 0))
        self._free.insert(i, (off, n))
        # coalesce with the neighbours
        if i + 1 < len(self._free) and off + n == self._free[i + 1][0]:
            self._free[i] = (off, n + self._free[i + 1][1])
            del self._free[i + 1]
        if i > 0 and self._free[i - 1][0] + self._free[i - 1][1] == off:
            self._free[i - 1] = (self._free[i - 1][0], self._free[i - 1][1] + self._free[i][1])
            del self._free[i]

    def view(self, off: int, nbytes: int) -> torch.Tensor:
        return self.buf[off:off + self._round(nbytes)]


def choose_store(full_bytes: int, ckpt_bytes: int, room: int, budget: int, policy: str = "auto",
                 full_share: int = 8) -> bool:
    """True: run the job with the full wavefield store; False: checkpointed.

    full_bytes / ckpt_bytes: the job's workspace in either mode; room: the
    largest block that can be placed now; budget: the arena's size.  The
    full store skips the recompute pass (two passes over the time axis
    instead of three) and gives identical results.  "auto" takes it when the
    full-store workspace is at most 1/full_share of the budget and can be
    placed now, so a few large jobs cannot crowd out concurrency; "full"
    whenever it can be placed now; "checkpoint" never."""
    if policy not in ("auto", "full", "checkpoint"):
        raise ValueError(f"store policy {policy!r}: 'auto', 'full' or 'checkpoint'")
    if policy == "checkpoint":
        return False
    if full_bytes <= ckpt_bytes:  # checkpointing would save nothing
        return True
    fits = full_bytes <= room
    if policy == "full":
        return fits
    return fits and full_bytes * full_share <= budget


def check_models(params: FWIParams, vel: torch.Tensor) -> None:
    """Refuse models the solver cannot run meaningfully: any non-finite
    value (the reference stops on NaN, pyacofwi2D.pyx:16-18 and
    geopvi/forward/fwi2d/posterior.py:159-160) or a velocity above the
    explicit scheme's stability limit (FWIParams.max_stable_velocity), where
    the fields grow without bound.  One reduction and one small copy on the
    caller's stream."""
    vmax = float(vel.abs().amax())   # NaN and Inf propagate through amax
    if not math.isfinite(vmax):
        raise ValueError("NaN or Inf in the velocity models")
    limit = params.max_stable_velocity()
    if vmax > limit:
        raise ValueError(f"velocity {vmax:.1f} m/s exceeds the stability limit {limit:.1f} m/s of dt={params.dt}, "
                         f"dx={params.dx}, dz={params.dz} (v*dt/dx = {vmax * params.dt / params.dx:.3f})")


@dataclass
class _Split:
    """A request whose samples run as several engine jobs (chunks of samples
    whose workspace fits the arena); they write into row slices of the
    request's outputs (allocated when the first chunk is admitted), and the
    request resolves when the last chunk finishes."""
    nsample: int
    remaining: int
    submitted: float
    admitted: float = 0.0
    ckpts: list = field(default_factory=list)
    syn: Optional[torch.Tensor] = None
    grad: Optional[torch.Tensor] = None
    misfit: Optional[torch.Tensor] = None
    outputs_ready: Optional[torch.cuda.Event] = None   # recorded after the outputs are zeroed
    lane: int = 0                                       # the lane whose stream made the outputs
    lanes: set = field(default_factory=set)             # lanes its chunks ran on
    for_memory: bool = False                            # split because it exceeds half the arena


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
    split: Optional[_Split] = None   # set on the chunks of a split request
    rows: Tuple[int, int] = (0, 0)   # this chunk's samples in the request


class FWIServer:
    """Serve FWI-2D misfit+gradient requests with continuous batching.

    mem_budget: bytes of the workspace arena (default: 80 % of the free
    memory when the server starts, counting blocks torch's allocator holds
    cached; the arena goes back when the server closes).  The rest holds
    the requests' inputs and outputs.  A request whose checkpointed
    workspace exceeds half the arena is split into chunks of samples that
    fit, run as separate jobs into row slices of its outputs.  steps_per_tick:
    time steps advanced between admission/completion checks (rounded to
    whole graph segments by the engine).  store_policy: how a request
    without an explicit ckpt_interval keeps its forward wavefield (see
    choose_store).  validate_models: refuse, at submit(), models with
    non-finite values or velocities above the scheme's stability limit
    (check_models).  log_path: append one JSON line per finished or failed
    request (ids, times, latency, sizes, store mode).  A job whose misfit
    comes out non-finite (for instance from NaN or Inf in its observed data)
    fails alone with FloatingPointError; the jobs beside it are unaffected.
    lanes: engines, each on its own stream (1 or 2); a request joins the
    lane with the fewest units in flight.  With two, one lane's step
    launches fill the GPU while the other's drain.  A job's results are the
    same bits on either lane.  split_lanes: with two lanes, run a request of
    two or more samples as two chunks of samples (one per lane), so both
    lanes work on it at once.
    """

    def __init__(self, device: int = 0, mem_budget: Optional[int] = None, steps_per_tick: int = 64,
                 max_units: int = 16384, max_tiles: int = 1 << 21, store_policy: str = "auto",
                 validate_models: bool = True, log_path: Optional[str] = None, lanes: int = 2,
                 split_lanes: bool = True):
        choose_store(0, 0, 0, 1, store_policy)  # validates the name
        if lanes not in (1, 2):
            raise ValueError(f"lanes={lanes}: 1 or 2")
        self.split_lanes = bool(split_lanes)
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        # The side stream first: with few hardware queues the runtime gives a
        # new stream a least-used queue, and two lane streams made back to
        # back were seen to share one (then they run one after the other).
        self._side = torch.cuda.Stream(self.device)  # reads finished outputs without waiting for queued steps
        self.streams = [torch.cuda.Stream(self.device) for _ in range(lanes)]
        self.stream = self.streams[0]
        if mem_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            cached = torch.cuda.memory_reserved(self.device) - torch.cuda.memory_allocated(self.device)
            mem_budget = int((free + cached) * 0.8)
        self.arena = Arena(mem_budget, self.device)
        self.mem_budget = self.arena.size
        self.store_policy = store_policy
        self.validate_models = bool(validate_models)
        self.steps_per_tick = int(steps_per_tick)
        self._log = open(log_path, "a", buffering=1) if log_path else None
        self._next_req = 0
        self.engines = [Engine(st, max_units=max_units, max_tiles=max_tiles) for st in self.streams]
        self.max_units = max_units
        self._queue: collections.deque = collections.deque()
        self._inflight: dict = {}            # (lane, job id) -> request
        self._units_in_use = [0] * lanes     # per lane
        self._cv = threading.Condition()
        self._stop = False
        self._error: Optional[BaseException] = None
        self.completed: list = []
        self.arena_waits = 0  # requests (or chunks) whose admission waited for arena space
        self._thread = threading.Thread(target=self._run, name="fwi-server", daemon=True)
        self._thread.start()

    # -- client side -----------------------------------------------------------
    def submit(self, params: FWIParams, vel, record_obs, run_fwi: int = 1, ckpt_interval: Optional[int] = None
               ) -> Future:
        """Queue one request; returns a Future of FWIResult.

        vel: (nsample, nz*nx) models (device tensor kept as is, host arrays
        are copied); record_obs: (ns*nt*nr,) observed data on the device."""
        solver.validate(params)
        if params.laplace_slover == 1:
            raise ValueError("the server runs finite-difference jobs (laplace_slover=0); run pseudo-spectral ones "
                             "through solver.fwi_batch")
        t_sub = time.perf_counter()
        if not isinstance(vel, torch.Tensor) or not vel.is_cuda:
            vel = torch.as_tensor(np.asarray(vel, np.float32)).to(self.device)
        vel = vel.to(torch.float32).reshape(-1, params.ni).contiguous()
        if self.validate_models:
            check_models(params, vel)
        obs = record_obs if isinstance(record_obs, torch.Tensor) else torch.as_tensor(record_obs)
        obs = obs.to(self.device, torch.float32).reshape(-1).contiguous()
        if obs.numel() != params.record_len:
            raise ValueError(f"record_obs has {obs.numel()} values, expected {params.record_len}")
        # without an explicit interval the store mode is chosen at admission;
        # the checkpointed workspace (the smallest) must fit the arena
        ckpt = None if ckpt_interval is None else int(ckpt_interval)
        ck_plan = best_ckpt_interval(params) if ckpt is None else ckpt
        n = vel.shape[0]
        need = job_workspace_bytes(params, n, ck_plan)
        chunk = n
        for_memory = need > self.arena.size // 2
        if for_memory:
            # Too large for one job next to others: split the samples into
            # chunks whose workspace is at most half the arena (two chunks
            # can then run side by side), each written into its rows.
            chunk = self._chunk_samples(params, n, ck_plan, self.arena.size // 2)
            if chunk == 0:
                raise MemoryError(f"one sample needs {job_workspace_bytes(params, 1, ck_plan) / 1e9:.1f} GB of "
                                  f"workspace; the arena has {self.arena.size / 1e9:.1f} GB")
        elif self.split_lanes and len(self.engines) > 1 and n >= 2:
            # Two chunks, one per lane: both lanes work on the request at
            # once (the same rows, the same bits).
            chunk = (n + 1) // 2
        fut: Future = Future()
        reqs = []
        with self._cv:
            rid = self._next_req
            self._next_req += 1
        if chunk >= n:
            ev = torch.cuda.Event()
            ev.record()  # inputs written on the caller's stream are ready here
            reqs.append(_Request(params, vel, obs, int(run_fwi), ckpt, need, ev, fut, t_sub, extra={"id": rid}))
        else:
            split = _Split(n, (n + chunk - 1) // chunk, t_sub, for_memory=for_memory)
            ev = torch.cuda.Event()
            ev.record()  # inputs written on the caller's stream are ready here
            for lo in range(0, n, chunk):
                hi = min(n, lo + chunk)
                reqs.append(_Request(params, vel[lo:hi], obs, int(run_fwi), ckpt,
                                     job_workspace_bytes(params, hi - lo, ck_plan), ev, fut, t_sub,
                                     extra={"id": rid}, split=split, rows=(lo, hi)))
        with self._cv:
            if self._error is not None:
                raise RuntimeError("server stopped after an error") from self._error
            self._queue.extend(reqs)
            self._cv.notify()
        return fut

    @staticmethod
    def _chunk_samples(params: FWIParams, n: int, ckpt: int, limit: int) -> int:
        """Most samples (<= n) whose checkpointed workspace is <= limit."""
        lo, hi = 0, n
        while lo < hi:
            mid = (lo + hi + 1) // 2
            if job_workspace_bytes(params, mid, ckpt) <= limit:
                lo = mid
            else:
                hi = mid - 1
        return lo

    def close(self, wait: bool = True):
        with self._cv:
            self._stop = True
            self._cv.notify()
        if wait:
            self._thread.join()
        for e in self.engines:
            e.close()
        if self._log is not None:
            self._log.close()
            self._log = None
        if wait:
            self.arena = None
            torch.cuda.empty_cache()  # the arena's memory goes back to the device

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def stats(self) -> dict:
        lat = np.array([r.latency for r in self.completed])
        if lat.size == 0:
            return {"completed": 0}
        return {"completed": int(lat.size), "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
                "arena_waits": self.arena_waits}

    # -- server thread ---------------------------------------------------------
    def _place(self, req):
        """(ckpt, workspace bytes, arena offset) for the request, or None if
        its workspace cannot be placed now."""
        n = req.vel.shape[0]
        if req.ckpt is not None:
            need = job_workspace_bytes(req.params, n, req.ckpt)
            off = self.arena.alloc(need)
            return None if off is None else (req.ckpt, need, off)
        full = job_workspace_bytes(req.params, n, 0)
        if choose_store(full, req.need, self.arena.largest(), self.arena.size, self.store_policy):
            off = self.arena.alloc(full)
            if off is not None:
                return 0, full, off
        off = self.arena.alloc(req.need)
        return None if off is None else (best_ckpt_interval(req.params), req.need, off)

    def _lane_busy(self, lane: int) -> bool:
        return any(k[0] == lane for k in self._inflight)

    def _admit(self):
        while self._queue:
            req = self._queue[0]
            units = req.vel.shape[0] * req.params.ns
            lane = int(np.argmin(self._units_in_use))
            if self._lane_busy(lane) and self._units_in_use[lane] + units > self.max_units:
                return
            placed = self._place(req)
            if placed is None:
                # every job's checkpointed workspace fits the empty arena
                # (submit() checks), so running jobs will make room
                if not req.extra.get("waited"):
                    req.extra["waited"] = True
                    self.arena_waits += 1
                return
            ckpt, need, off = placed
            outputs = None
            sp = req.split
            st = self.streams[lane]
            try:
                # outputs are made (and the gradient zeroed) on the lane's stream
                with torch.cuda.stream(st):
                    if sp is not None:
                        if sp.syn is None:  # first chunk: the request's outputs
                            n, p = sp.nsample, req.params
                            sp.syn = torch.empty(n, p.record_len, dtype=torch.float32, device=self.device)
                            sp.grad = torch.zeros(n, p.ni, dtype=torch.float32, device=self.device)
                            sp.misfit = torch.empty(n, dtype=torch.float64, device=self.device)
                            sp.outputs_ready = torch.cuda.Event()
                            sp.outputs_ready.record(st)
                            sp.lane = lane
                        lo, hi = req.rows
                        outputs = (sp.syn[lo:hi], sp.grad[lo:hi], sp.misfit[lo:hi])
                    job = EngineJob(req.params, req.vel, req.obs, ckpt, req.run_fwi,
                                    workspace=self.arena.view(off, need), outputs=outputs)
            except torch.OutOfMemoryError as exc:  # the outputs (outside the arena)
                self.arena.release(off, need)
                if self._inflight:
                    return
                # nothing running will free memory
                self._fail_queued(req, MemoryError(f"no room for the outputs: {exc}"))
                continue
            st.wait_event(req.ready)
            if sp is not None and sp.lane != lane:
                st.wait_event(sp.outputs_ready)
            try:
                self.engines[lane].submit(job)
            except _lib.GpviError as exc:
                if exc.code == _lib.GPVI_ERR_HIP:
                    raise  # the device failed: stop, failing every waiter
                job.workspace = None
                self.arena.release(off, need)
                # The engine's unit slots (a contiguous run per job) or tile
                # list are full: wait for running jobs, or fail the request
                # when nothing runs.  Any other refusal is about this request
                # alone: fail it and go on.
                if exc.code == _lib.GPVI_ERR_CAPACITY and self._inflight:
                    return
                self._fail_queued(req, _lib.GpviError(exc.code, exc.msg))
                continue
            self._queue.popleft()
            req.ckpt, req.need, req.extra["arena_offset"] = ckpt, need, off
            req.job = job
            req.admitted = time.perf_counter()
            if req.split is not None and not req.split.admitted:
                req.split.admitted = req.admitted
            req.extra["lane"] = lane
            if sp is not None:
                sp.lanes.add(lane)
            self._inflight[(lane, job.id)] = req
            self._units_in_use[lane] += units

    def _fail_queued(self, req, exc):
        """Drop the request at the head of the queue (with all its queued
        chunks, if it was split) and fail its future with `exc`.  `exc` must
        be a fresh exception with no traceback or context: one raised here
        would keep this thread's frames, and with them an arena view, alive
        through the future it is stored in."""
        exc.__context__ = None
        exc.__traceback__ = None
        self._queue.popleft()
        if req.split is not None:
            self._queue = collections.deque(r for r in self._queue if r.split is not req.split)
        if not req.future.done():
            req.future.set_exception(exc)
            self._write_log(req, None, f"failed: {exc}")

    def _nonfinite_rows(self, misfit: torch.Tensor):
        """Rows of a finished job's misfit that are NaN or Inf.  The job's
        kernels are complete, so the copy runs on a side stream and does not
        wait for the steps queued behind them."""
        with torch.cuda.stream(self._side):
            m = misfit.to("cpu")
        return np.flatnonzero(~np.isfinite(m.numpy()))

    def _write_log(self, req, res: Optional[FWIResult], status: str):
        if self._log is None:
            return
        p = req.params
        rec = {"id": req.extra.get("id"), "status": status, "grid": [p.nz, p.nx], "ns": p.ns, "nt": p.nt,
               "nsample": req.split.nsample if req.split is not None else int(req.vel.shape[0]),
               "submitted": req.submitted}
        if res is not None:
            rec.update(admitted=res.admitted, finished=res.finished, latency_s=res.latency,
                       ckpt_interval=res.extra.get("ckpt_interval"), chunks=res.extra.get("chunks", 1))
        self._log.write(json.dumps(rec) + "\n")

    def _reap(self):
        done = [k for k in self._inflight if self.engines[k[0]].status(k[1]) == 1]
        for k in done:
            req = self._inflight.pop(k)
            self.engines[k[0]].release(k[1])
            self._units_in_use[k[0]] -= req.vel.shape[0] * req.params.ns
            job = req.job
            job.workspace = None
            # the job's last kernels are complete (status 1), so its arena
            # block can be handed to the next job
            self.arena.release(req.extra["arena_offset"], req.need)
            bad = self._nonfinite_rows(job.misfit)
            sp = req.split
            if len(bad) and not req.future.done():
                lo = req.rows[0] if sp is not None else 0
                exc = FloatingPointError(f"non-finite misfit for sample(s) {(bad + lo).tolist()}: unstable model "
                                         "or NaN/Inf in the observed data")
                req.future.set_exception(exc)
                self._write_log(req, None, f"failed: {exc}")
                if sp is not None:  # drop the request's chunks still queued
                    with self._cv:
                        self._queue = collections.deque(r for r in self._queue if r.split is not sp)
            if sp is None:
                if req.future.done():
                    continue
                res = FWIResult(job.misfit, job.grad, job.syn, req.submitted, req.admitted, time.perf_counter())
                res.extra["ckpt_interval"] = req.ckpt
                res.extra["lane"] = req.extra["lane"]
                res.extra["lanes"] = [req.extra["lane"]]
            else:
                sp.remaining -= 1
                sp.ckpts.append(req.ckpt)
                if sp.remaining or req.future.done():
                    continue
                res = FWIResult(sp.misfit, sp.grad, sp.syn, sp.submitted, sp.admitted, time.perf_counter())
                res.extra["ckpt_interval"] = sp.ckpts[0] if len(set(sp.ckpts)) == 1 else list(sp.ckpts)
                res.extra["chunks"] = len(sp.ckpts)
                res.extra["lane"] = sp.lane   # the first chunk's
                res.extra["lanes"] = sorted(sp.lanes)
                res.extra["split_for_memory"] = sp.for_memory
            self.completed.append(res)
            req.future.set_result(res)
            self._write_log(req, res, "ok")

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
                        busy = [lane for lane in range(len(self.engines)) if self._lane_busy(lane)]
                        for lane in busy:
                            self.engines[lane].advance(self.steps_per_tick)
                        n_before = len(self._inflight)
                        self._reap()
                        if len(self._inflight) == n_before and all(
                                self.engines[lane].pending_steps() == 0 for lane in busy):
                            # every active job has all its steps queued: wait
                            # for the device (one lane) instead of polling the
                            # events, or poll every 0.5 ms (two lanes: either
                            # may finish first)
                            if len(busy) == 1:
                                self.engines[busy[0]].sync()
                            else:
                                time.sleep(5e-4)
        except BaseException as exc:  # surface the failure to every waiter
            self._error = exc
            for req in list(self._queue) + list(self._inflight.values()):
                if not req.future.done():
                    req.future.set_exception(exc)
            raise
