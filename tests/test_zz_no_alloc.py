"""gpvi_fwi2d_batch_static: the batched solver on the library's fixed device
scratch (no workspace argument, nothing allocated), called through the C
ABI and held to gpvi_fwi2d_batch on the same inputs, bit for bit, on two
calls in a row, the second with a smaller problem; a job larger than the
scratch and a pseudo-spectral job are refused with their status codes."""
import ctypes

import numpy as np
import pytest
import torch

from geopvi_amd import _lib, synth
from geopvi_amd.forward.fwi2d import solver

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)
SCRATCH = 64 << 20   # GPVI_STATIC_SCRATCH_BYTES


def inputs(name, nsample, seed, **over):
    p = synth.workload_params(name, **over)
    vt = synth.gradient_anomaly_model(p.nz, p.nx, p.dx, p.dz, seed=seed)
    vel = synth.gaussian_samples(synth.smooth(vt, 2.0), 30.0, nsample, seed + 1).astype(np.float32)
    obs = solver.forward_batch(p, torch.from_numpy(vt.astype(np.float32).reshape(1, -1)).to(DEV))[0].contiguous()
    return p, torch.from_numpy(vel).to(DEV), obs


def outputs(p, n):
    return (torch.full((n, p.record_len), 7.0, device=DEV), torch.zeros(n, p.ni, device=DEV),
            torch.full((n,), -1.0, dtype=torch.float64, device=DEV))


def static_call(p, vel, obs, ckpt=0, run_fwi=1):
    n = vel.shape[0]
    syn, grad, mis = outputs(p, n)
    c = p.to_c()
    rc = _lib.load().gpvi_fwi2d_batch_static(ctypes.byref(c), n, vel.data_ptr(), obs.data_ptr(), syn.data_ptr(),
                                              grad.data_ptr(), mis.data_ptr(), run_fwi, ckpt,
                                              ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc, syn, grad, mis


def reference_call(p, vel, obs, ckpt=0, run_fwi=1):
    syn, grad, mis = solver.fwi_batch(p, vel, obs, run_fwi=run_fwi, ckpt_interval=ckpt)
    return syn, grad, mis


def test_two_calls_match_the_workspace_entry_point():
    big = inputs("small", 4, 41, nt_interval=3)               # the larger call first
    small = inputs("tiny", 2, 42)
    assert solver.batch_workspace_bytes(big[0], 4, 0) <= SCRATCH
    for (p, vel, obs), ckpt in ((big, 0), (small, 64)):
        rc, syn, grad, mis = static_call(p, vel, obs, ckpt)
        assert rc == 0, _lib.last_error()
        ref = reference_call(p, vel, obs, ckpt)
        torch.cuda.synchronize()
        print(f"{p.nz}x{p.nx} x{vel.shape[0]} ckpt {ckpt}: misfit {mis.cpu().numpy()}", flush=True)
        assert torch.equal(syn, ref[0]) and torch.equal(grad, ref[1]) and torch.equal(mis, ref[2])


def test_refuses_jobs_beyond_the_scratch_and_pseudo_spectral_jobs():
    p, vel, obs = inputs("c0", 1, 43)
    assert solver.batch_workspace_bytes(p, 1, 0) > SCRATCH    # full store of one 120x200 sample: too large
    rc, syn, _, _ = static_call(p, vel, obs, 0)
    assert rc == -6 and "scratch" in _lib.last_error()          # GPVI_ERR_WORKSPACE
    assert bool((syn == 7.0).all())                               # nothing was written
    q = synth.workload_params("tiny", laplace_slover=1)
    vq = torch.full((1, q.ni), 2000.0, device=DEV)
    rc, _, _, _ = static_call(q, vq, torch.zeros(q.record_len, device=DEV))
    assert rc == -2                                               # GPVI_ERR_UNSUPPORTED
