"""Likelihood adapter of 2-D FWI, drop-in for
geopvi/forward/fwi2d/posterior.py (``ForwardModel``,
``prepare_fwi_parameters``, ``Posterior``).

Same constructor and method signatures and the same outputs.  What changes
is underneath: ``fwi``/``fwi_parallel`` evaluate all samples x all shots in
one batched GPU solve instead of a per-call process pool (so
``num_processes`` is accepted and ignored), the observed data stay resident
on the GPU, and ``log_prob`` keeps CUDA inputs on the device end to end.
"""
from __future__ import annotations

from typing import Optional, Union

import numpy as np
import torch
from torch.autograd import Function

from . import aco2d, solver
from .params import FWIParams, prepare_fwi_parameters

__all__ = ["ForwardModel", "prepare_fwi_parameters", "Posterior"]


class ForwardModel(Function):
    """Wraps a host function ``func(x) -> (loss (m,), grad (m, ndim))`` as an
    autograd node: forward returns the loss, backward scales the stored
    gradient by the incoming gradient (posterior.py:8-24)."""

    @staticmethod
    def forward(ctx, input, func):
        loss, grad = func(input)
        ctx.save_for_backward(input, torch.as_tensor(grad))
        return torch.as_tensor(loss)

    @staticmethod
    def backward(ctx, grad_output):
        _, grad = ctx.saved_tensors
        return grad_output[..., None] * grad, None


class Posterior:
    """Unnormalised log-posterior of 2-D acoustic FWI.

    data: observed records, flat (ns*nt*nr,) float32 in [ns][nt][nr] order
    config: configparser object with an [FWI] section (the example config.ini)
    vel_fixed: velocity of the fixed (mask == False) cells
    sigma: data noise; logp = -misfit/sigma**2 + log_prior(x)
    mask: boolean (nz*nx,) of the free cells (default: all free)
    paramfile: where the solver's parameter file is written

    Extra keyword-only arguments (defaults keep the reference behaviour):
    device: GPU to run on (default: the current one)
    ckpt_interval: 0 = keep every nt_interval-th wavefield (the reference's
        store); > 0 = checkpoint every this many steps and recompute;
        "auto" = per call, the full store when its workspace takes at most
        3/4 of the device memory free to this process (torch's cached
        blocks counted), else the interval that minimises memory (the
        results are the same bits either way; the full store saves the
        recompute pass)
    max_batch: most samples per batched solve (default: all at once)
    """

    def __init__(self, data, config, vel_fixed=1950, sigma=0.1, num_processes=1, log_prior=None, mask=None,
                 paramfile="input_params.txt", *, device=None, ckpt_interval: Union[int, str] = 0,
                 max_batch: Optional[int] = None):
        self.data = data
        self.log_prior = log_prior
        self.num_processes = num_processes
        self.sigma = sigma
        self.vel_water = vel_fixed
        self.paramfile = paramfile
        if mask is None:
            nx = config.getint("FWI", "nx")
            nz = config.getint("FWI", "nz")
            mask = np.full((nx * nz), True)
        self.mask = mask
        prepare_fwi_parameters(self.paramfile, config)
        self.params = FWIParams.read(self.paramfile)
        solver.validate(self.params)
        if not torch.cuda.is_available():
            raise RuntimeError("geopvi_amd: no HIP device visible (the FWI-2D solver runs on the GPU only)")
        self.device = torch.device(device) if device is not None else torch.device(
            "cuda", torch.cuda.current_device())
        if ckpt_interval != "auto":
            ckpt_interval = int(ckpt_interval)
        self.ckpt_interval = ckpt_interval
        self.last_ckpt_interval: Optional[int] = None   # the interval the last solve used
        self.max_batch = max_batch
        self._data_dev = torch.as_tensor(np.ascontiguousarray(data, dtype=np.float32).reshape(-1),
                                         device=self.device)
        self._mask_idx = torch.as_tensor(np.flatnonzero(np.asarray(mask).reshape(-1)), device=self.device)

    # -- reference methods ---------------------------------------------------
    def solver(self, vel):
        """One model through the plain solver: (loss, -grad[mask])."""
        data_syn, grad = aco2d.fwi(vel, self.data, paramfile=self.paramfile)
        grad = grad[self.mask]
        loss = 0.5 * np.sum((data_syn - self.data) ** 2)
        return loss, -grad

    def _full_models(self, x) -> np.ndarray:
        if not isinstance(x, np.ndarray):
            x = x.detach().cpu().numpy().astype(np.float64)
        m = x.shape[0]
        vel = np.full([m, self.mask.size], self.vel_water, dtype=np.float64)
        vel[:, self.mask] = x
        return vel

    def fwi(self, x):
        """Misfit (m,) and misfit gradient (m, ndim) of every sample, batched."""
        vel = self._full_models(x)
        if np.isnan(vel).any():
            raise ValueError("NaN occured in sample!")
        m = vel.shape[0]
        loss = np.zeros(m)
        grad = np.zeros((m, int(np.count_nonzero(self.mask))))
        v32 = torch.from_numpy(np.float32(vel)).to(self.device)
        step = self.max_batch or m
        for lo in range(0, m, step):
            hi = min(m, lo + step)
            _, g, mis = solver.fwi_batch(self.params, v32[lo:hi].contiguous(), self._data_dev,
                                         ckpt_interval=self._ckpt(hi - lo))
            loss[lo:hi] = mis.cpu().numpy()
            grad[lo:hi] = -(g.cpu().numpy()[:, self.mask])
        return loss, grad

    def fwi_parallel(self, x):
        """Same as ``fwi``: the batched solve replaces the process pool."""
        return self.fwi(x)

    def log_prob(self, x):
        """log-posterior of each row of x (nsample, ndim)."""
        if torch.isnan(x).any():
            raise ValueError("NaN occured in sample!")
        if x.is_cuda:
            loss = self._misfit_device(x)
        else:
            loss = ForwardModel.apply(x, self.fwi)
        logp = -loss / self.sigma ** 2 + self.log_prior(x)
        return logp

    def _ckpt(self, nsample: int) -> int:
        """The checkpoint interval of a solve of ``nsample`` models."""
        if self.ckpt_interval != "auto":
            ck = self.ckpt_interval
        else:
            ck = solver.auto_ckpt_interval(self.params, nsample, self.device)
        self.last_ckpt_interval = ck
        return ck

    # -- device-resident path ------------------------------------------------
    def _misfit_device(self, x: torch.Tensor) -> torch.Tensor:
        m = x.shape[0]
        vel = torch.full((m, self.mask.size), float(self.vel_water), dtype=x.dtype, device=x.device)
        vel[:, self._mask_idx.to(x.device)] = x
        outs = []
        step = self.max_batch or m
        for lo in range(0, m, step):
            chunk = vel[lo:lo + step]
            outs.append(solver.fwi2d_misfit(chunk, self._data_dev.to(x.device), self.params,
                                            self._ckpt(chunk.shape[0])))
        return torch.cat(outs) if len(outs) > 1 else outs[0]
