"""Device-level entry points of the FWI-2D solver (torch tensors on the GPU).

Every function here hands device pointers and torch's current HIP stream to
the C ABI of libgeopvi_fwi2d.so; the work is enqueued asynchronously on that
stream.  Workspaces are torch allocations, so torch's caching allocator
recycles them across calls.

Sign convention (geopvi/forward/fwi2d/posterior.py:108-110): the solver's
``grad`` output is the reference's imaging-condition gradient; the gradient
of the misfit with respect to velocity is ``-grad``.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import torch

from ... import _lib
from .params import FWIParams


def _stream(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _dev(t: torch.Tensor, name: str, numel: Optional[int] = None) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name} must be a CUDA (HIP) tensor")
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")
    return t


def _params_ptr(params: FWIParams):
    c = params.to_c()
    return c, ctypes.byref(c)


def validate(params: FWIParams) -> None:
    c, pc = _params_ptr(params)
    _lib.check(_lib.load().gpvi_fwi2d_validate(pc))


def _workspace(nbytes: int, device: torch.device) -> torch.Tensor:
    ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)
    if ws.data_ptr() % 256:
        raise RuntimeError("torch returned a workspace that is not 256-byte aligned")
    return ws


def plain_workspace_bytes(params: FWIParams) -> int:
    c, pc = _params_ptr(params)
    n = _lib.load().gpvi_fwi2d_plain_workspace_bytes(pc)
    if n == 0:
        validate(params)
    return int(n)


def batch_workspace_bytes(params: FWIParams, nsample: int, ckpt_interval: int = 0) -> int:
    c, pc = _params_ptr(params)
    n = _lib.load().gpvi_fwi2d_batch_workspace_bytes(pc, int(nsample), int(ckpt_interval))
    if n == 0:
        validate(params)
        raise ValueError(f"nsample={nsample}")
    return int(n)


def model_workspace_bytes(params: FWIParams, nsample: int) -> int:
    """Workspace of forward modelling only (fields and records, no store)."""
    c, pc = _params_ptr(params)
    n = _lib.load().gpvi_fwi2d_model_workspace_bytes(pc, int(nsample))
    if n == 0:
        validate(params)
        raise ValueError(f"nsample={nsample}")
    return int(n)


def schedule(params: FWIParams, ckpt_interval: int = 0, steps_per_launch: int = 1) -> dict:
    """Adjoint schedule and launch counts of a batched evaluation.

    Mirrors the library's schedule: ckpt_interval <= 0 or >= nt keeps the
    full store (one segment); otherwise it is rounded up to a multiple of
    nt_interval and the backward pass recomputes each segment from its
    checkpoint before running its adjoint.  Each launch advances every unit
    by `steps_per_launch` steps of one pass (a pass of odd length ends with
    a one-step launch)."""
    nt, k, tb = params.nt, params.nt_interval, int(steps_per_launch)
    cdiv = lambda a, b: -(-a // b)  # noqa: E731
    c = cdiv(ckpt_interval, k) * k if 0 < ckpt_interval < nt else nt
    nseg = cdiv(nt, c)
    if nseg <= 1:
        nseg, c = 1, nt
    last = nt - (nseg - 1) * c
    fwd = cdiv(nt, tb)
    if nseg == 1:
        back = cdiv(nt, tb)
    else:
        back = (nseg - 1) * 2 * cdiv(c, tb) + 2 * cdiv(last, tb)
    return {"nseg": nseg, "ckpt": c, "last_len": last, "launches_forward": fwd, "launches_backward": back,
            "launches": fwd + back, "recomputed_steps": 0 if nseg == 1 else nt}


# ---------------------------------------------------------------------------
# Plain solver: one sample, shots in sequence, one launch per time step.

def fwi_plain(params: FWIParams, vel: torch.Tensor, record_obs: torch.Tensor, run_fwi: int = 1,
              grad: Optional[torch.Tensor] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """fwi_2D on device memory (fdtd2D_modelling.c:18-353).

    vel (nz*nx,) float32, record_obs (ns*nt*nr,) float32 -> (record_syn,
    grad); ``grad`` is accumulated into if given (as the reference does),
    else starts from zero."""
    _dev(vel, "vel", params.ni)
    _dev(record_obs, "record_obs", params.record_len)
    dev = vel.device
    syn = torch.empty(params.record_len, dtype=torch.float32, device=dev)
    if grad is None:
        grad = torch.zeros(params.ni, dtype=torch.float32, device=dev)
    else:
        _dev(grad, "grad", params.ni)
    ws = _workspace(plain_workspace_bytes(params), dev)
    c, pc = _params_ptr(params)
    _lib.check(_lib.load().gpvi_fwi2d_fwi_device(pc, vel.data_ptr(), syn.data_ptr(), record_obs.data_ptr(),
                                                 grad.data_ptr(), int(run_fwi), ws.data_ptr(), ws.numel(),
                                                 _stream(dev)))
    return syn, grad


def forward_plain(params: FWIParams, vel: torch.Tensor) -> torch.Tensor:
    """forward_2D on device memory (fdtd2D_modelling.c:356-675)."""
    _dev(vel, "vel", params.ni)
    dev = vel.device
    syn = torch.empty(params.record_len, dtype=torch.float32, device=dev)
    ws = _workspace(plain_workspace_bytes(params), dev)
    c, pc = _params_ptr(params)
    _lib.check(_lib.load().gpvi_fwi2d_forward_device(pc, vel.data_ptr(), syn.data_ptr(), ws.data_ptr(),
                                                     ws.numel(), _stream(dev)))
    return syn


# ---------------------------------------------------------------------------
# Batched solver: every (sample, shot) unit in each step launch.

def _batch_call(params, nsample, vel, obs, syn, grad, misfit, run_fwi, ckpt, phase, flags, ws, dev):
    c, pc = _params_ptr(params)
    ptr = lambda t: t.data_ptr() if t is not None else None  # noqa: E731
    _lib.check(_lib.load().gpvi_fwi2d_batch(
        pc, int(nsample), ptr(vel), ptr(obs), ptr(syn), ptr(grad), ptr(misfit), int(run_fwi), int(ckpt),
        int(phase), int(flags), ws.data_ptr(), ws.numel(), _stream(dev)))


def auto_ckpt_interval(params: FWIParams, nsample: int, device) -> int:
    """Checkpoint interval of a batched solve of ``nsample`` models: 0 (the
    full wavefield store) when its workspace fits in 3/4 of the device
    memory that is free or cached by torch, else the interval that minimises
    a unit's memory (``engine.best_ckpt_interval``)."""
    free, _ = torch.cuda.mem_get_info(device)
    free += torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
    if batch_workspace_bytes(params, nsample, 0) <= 0.75 * free:
        return 0
    from ...engine import best_ckpt_interval
    return best_ckpt_interval(params)


def fwi_batch(params: FWIParams, vel: torch.Tensor, record_obs: torch.Tensor, run_fwi: int = 1,
              ckpt_interval: int = 0, graph: bool = True,
              workspace: Optional[torch.Tensor] = None):
    """Misfit and gradient of ``nsample`` models in one batched evaluation.

    vel (nsample, nz*nx) float32, record_obs (ns*nt*nr,) shared by every
    sample.  Returns (record_syn (nsample, ns*nt*nr), grad (nsample, nz*nx),
    misfit (nsample,) float64 = 0.5*sum((syn-obs)^2))."""
    if vel.dim() == 1:
        vel = vel.view(1, -1)
    nsample = vel.shape[0]
    _dev(vel, "vel", nsample * params.ni)
    _dev(record_obs, "record_obs", params.record_len)
    dev = vel.device
    syn = torch.empty(nsample, params.record_len, dtype=torch.float32, device=dev)
    grad = torch.zeros(nsample, params.ni, dtype=torch.float32, device=dev)
    misfit = torch.empty(nsample, dtype=torch.float64, device=dev)
    need = batch_workspace_bytes(params, nsample, ckpt_interval)
    if workspace is None or workspace.numel() < need:
        workspace = _workspace(need, dev)
    flags = 0 if graph else _lib.GPVI_FLAG_NO_GRAPH
    _batch_call(params, nsample, vel, record_obs, syn, grad, misfit, run_fwi, ckpt_interval,
                _lib.GPVI_PHASE_ALL, flags, workspace, dev)
    return syn, grad, misfit


def forward_batch(params: FWIParams, vel: torch.Tensor, graph: bool = True) -> torch.Tensor:
    """Records of ``nsample`` models (forward modelling only, no wavefield store)."""
    if vel.dim() == 1:
        vel = vel.view(1, -1)
    nsample = vel.shape[0]
    _dev(vel, "vel", nsample * params.ni)
    dev = vel.device
    syn = torch.empty(nsample, params.record_len, dtype=torch.float32, device=dev)
    dummy_obs = torch.zeros(params.record_len, dtype=torch.float32, device=dev)
    ws = _workspace(model_workspace_bytes(params, nsample), dev)
    flags = _lib.GPVI_FLAG_NO_STORE | (0 if graph else _lib.GPVI_FLAG_NO_GRAPH)
    _batch_call(params, nsample, vel, dummy_obs, syn, None, None, 1, 0, _lib.GPVI_PHASE_FORWARD, flags, ws, dev)
    return syn


# ---------------------------------------------------------------------------
# Autograd: forward pass kernel in forward(), hand-written adjoint in backward().

class FWI2DMisfit(torch.autograd.Function):
    """misfit = 0.5*||d_syn(vel) - d_obs||^2 per sample.

    forward() runs the forward propagation only (records, misfit, and either
    the full wavefield store or (p0, p1) checkpoints every ``ckpt_interval``
    steps, kept in a workspace saved on ctx); backward() runs the adjoint
    kernel, recomputing each checkpoint segment, and returns
    ``grad_output[:, None] * (-grad)``, the reference's gradient convention
    (geopvi/forward/fwi2d/posterior.py:16-24, :108-110)."""

    @staticmethod
    def forward(ctx, vel, record_obs, params: FWIParams, ckpt_interval: int = 0, run_fwi: int = 1):
        squeeze = vel.dim() == 1
        v32 = vel.detach().reshape(-1, params.ni).to(torch.float32).contiguous()
        obs = record_obs.detach().to(torch.float32).contiguous()
        nsample = v32.shape[0]
        _dev(v32, "vel", nsample * params.ni)
        _dev(obs, "record_obs", params.record_len)
        dev = v32.device
        syn = torch.empty(nsample, params.record_len, dtype=torch.float32, device=dev)
        misfit = torch.empty(nsample, dtype=torch.float64, device=dev)
        ws = _workspace(batch_workspace_bytes(params, nsample, ckpt_interval), dev)
        _batch_call(params, nsample, v32, obs, syn, None, misfit, run_fwi, ckpt_interval,
                    _lib.GPVI_PHASE_FORWARD, 0, ws, dev)
        ctx.params = params
        ctx.ckpt = int(ckpt_interval)
        ctx.run_fwi = int(run_fwi)
        ctx.squeeze = squeeze
        ctx.in_dtype = vel.dtype
        ctx.save_for_backward(ws, syn, obs)
        ctx.mark_non_differentiable(syn)
        out = misfit.to(vel.dtype) if vel.dtype.is_floating_point else misfit
        return (out[0] if squeeze else out), syn

    @staticmethod
    def backward(ctx, grad_misfit, grad_syn):
        ws, syn, obs = ctx.saved_tensors
        params = ctx.params
        nsample = syn.shape[0]
        grad = torch.zeros(nsample, params.ni, dtype=torch.float32, device=syn.device)
        _batch_call(params, nsample, None, obs, syn, grad, None, ctx.run_fwi, ctx.ckpt,
                    _lib.GPVI_PHASE_ADJOINT, 0, ws, syn.device)
        dmis = (-grad).to(ctx.in_dtype)
        gm = grad_misfit.reshape(-1, 1).to(ctx.in_dtype)
        gin = gm * dmis
        if ctx.squeeze:
            gin = gin.reshape(-1)
        return gin, None, None, None, None


def fwi2d_misfit(vel: torch.Tensor, record_obs: torch.Tensor, params: FWIParams, ckpt_interval: int = 0,
                 run_fwi: int = 1) -> torch.Tensor:
    """Differentiable per-sample misfit; see FWI2DMisfit."""
    misfit, _ = FWI2DMisfit.apply(vel, record_obs, params, ckpt_interval, run_fwi)
    return misfit
