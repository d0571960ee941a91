"""Python handle on the C-ABI continuous-batching engine (gpvi_engine_*).

The engine advances every active job by one time step per launch: all
(sample, shot) units of all jobs, of any grid size, share each launch
(replayed as graph segments).  Jobs join at a step boundary and leave when
their last step is done; their finalising kernels (shot-ordered gradient
sum, misfit) are queued right behind.  Inputs, outputs and the per-job
workspace are caller-owned device memory (torch tensors here).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _lib
from .forward.fwi2d import solver
from .forward.fwi2d.params import FWIParams


class EngineJob:
    """Device buffers of one submitted job; kept alive until it finishes."""

    def __init__(self, params: FWIParams, vel: torch.Tensor, obs: torch.Tensor, ckpt: int, run_fwi: int,
                 workspace: Optional[torch.Tensor] = None, outputs=None):
        """workspace: a uint8 device tensor of at least job_workspace_bytes
        bytes, 256-byte aligned (e.g. a slice of a server's arena); allocated
        here when None.  outputs: (record_syn (n, ns*nt*nr) float32,
        grad (n, nz*nx) float32, zeroed -- the solver adds to it --, misfit
        (n,) float64), contiguous device tensors (e.g. row slices of a larger
        request's outputs); allocated here when None."""
        self.params = params
        self.vel = vel
        self.obs = obs
        self.ckpt = ckpt
        self.run_fwi = run_fwi
        n = vel.shape[0]
        dev = vel.device
        if outputs is None:
            self.syn = torch.empty(n, params.record_len, dtype=torch.float32, device=dev)
            self.grad = torch.zeros(n, params.ni, dtype=torch.float32, device=dev)
            self.misfit = torch.empty(n, dtype=torch.float64, device=dev)
        else:
            self.syn, self.grad, self.misfit = outputs
            for t, shape, dt in ((self.syn, (n, params.record_len), torch.float32),
                                 (self.grad, (n, params.ni), torch.float32), (self.misfit, (n,), torch.float64)):
                if tuple(t.shape) != shape or t.dtype != dt or not t.is_contiguous() or t.device != dev:
                    raise ValueError(f"output {tuple(t.shape)} {t.dtype}: expected {shape} {dt}, contiguous on {dev}")
        need = job_workspace_bytes(params, n, ckpt)
        if workspace is None:
            workspace = torch.empty(need, dtype=torch.uint8, device=dev)
        elif workspace.numel() < need or workspace.data_ptr() % 256:
            raise ValueError(f"workspace: {workspace.numel()} bytes at {workspace.data_ptr():#x}, "
                             f"need {need} bytes 256-byte aligned")
        self.workspace = workspace
        self.id: Optional[int] = None

    @property
    def nbytes(self) -> int:
        return (self.workspace.numel() + self.syn.numel() * 4 + self.grad.numel() * 4 + self.misfit.numel() * 8)


def job_workspace_bytes(params: FWIParams, nsample: int, ckpt: int) -> int:
    return solver.batch_workspace_bytes(params, nsample, ckpt)


def job_bytes(params: FWIParams, nsample: int, ckpt: int) -> int:
    """Device memory one job holds while in flight: workspace + outputs."""
    return (job_workspace_bytes(params, nsample, ckpt) + nsample * params.record_len * 4 + nsample * params.ni * 4
            + nsample * 8)


def best_ckpt_interval(params: FWIParams) -> int:
    """Checkpoint interval (multiple of nt_interval) that minimises a unit's
    memory: nt/C checkpoint pairs of padded fields against C/k stored inner
    fields, C* = sqrt(2 * nt * k * ntp / ni)."""
    k = params.nt_interval
    c = (2.0 * params.nt * k * params.ntp / params.ni) ** 0.5
    c = max(k, int(round(c / k)) * k)
    return c if c < params.nt else 0


class Engine:
    def __init__(self, stream: torch.cuda.Stream, max_units: int = 8192, max_tiles: int = 1 << 20):
        self.stream = stream
        h = ctypes.c_void_p()
        _lib.check(_lib.load().gpvi_engine_create(ctypes.c_void_p(stream.cuda_stream), int(max_units),
                                                  int(max_tiles), ctypes.byref(h)))
        self._h = h
        self.max_units = max_units

    def close(self):
        if self._h:
            _lib.load().gpvi_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, job: EngineJob) -> int:
        p = job.params.to_c()
        jid = ctypes.c_longlong()
        _lib.check(_lib.load().gpvi_engine_submit(
            self._h, ctypes.byref(p), int(job.vel.shape[0]), job.vel.data_ptr(), job.obs.data_ptr(),
            job.syn.data_ptr(), job.grad.data_ptr(), job.misfit.data_ptr(), int(job.run_fwi), int(job.ckpt),
            job.workspace.data_ptr(), job.workspace.numel(), ctypes.byref(jid)))
        job.id = jid.value
        return job.id

    def advance(self, nsteps: int) -> int:
        return _lib.check(_lib.load().gpvi_engine_advance(self._h, int(nsteps)))

    def status(self, job_id: int) -> int:
        return _lib.check(_lib.load().gpvi_engine_job_status(self._h, int(job_id)))

    def release(self, job_id: int) -> None:
        _lib.check(_lib.load().gpvi_engine_release(self._h, int(job_id)))

    def cancel(self, job_id: int) -> None:
        _lib.check(_lib.load().gpvi_engine_cancel(self._h, int(job_id)))

    def active_jobs(self) -> int:
        return _lib.load().gpvi_engine_active_jobs(self._h)

    def active_units(self) -> int:
        return _lib.load().gpvi_engine_active_units(self._h)

    def pending_steps(self) -> int:
        return _lib.load().gpvi_engine_pending_steps(self._h)

    def sync(self) -> None:
        _lib.check(_lib.load().gpvi_engine_sync(self._h))
