"""The reference's is_fwi = 0 mode (pyacofwi2D.pyx:11-28, fdtd2D_modelling.c
:274-277, 920-924, 972-974) and the adjoint's injection terms on the GPU.

run_fwi = 0: the residual becomes the observed data itself, the wavefield
store keeps the field p2 instead of its second time derivative, and the
gradient accumulates store * p2 * nt_interval * dt in float.  Every entry
point that takes run_fwi is held to the oracle: the plain solver, the
batched solver with and without checkpoints, the autograd Function, the
host-pointer fwi_2D and aco2d.fwi(is_fwi=0).

The point-source / point-receiver cases of the oracle's closed-form pins
(tests/test_oracle_adjoint_pins.py: residual = time-reversed wavelet, so the
adjoint field is the receiver's forward field reversed) run here through the
batched solver for both modes and k = 1, 2, against the oracle.

The adjoint phase of the autograd Function keeps the forward phase's launch
layout even when the tuning changes between forward() and backward().
"""
import ctypes

import numpy as np
import pytest
import torch

from helpers import GRAD_RTOL, MISFIT_RTOL, rel_err, rel_l2
from geopvi_amd import _lib, synth
from geopvi_amd.forward.fwi2d import aco2d, solver
from test_oracle_adjoint_pins import A, B, hetero_model, point_params, reversed_residual_obs

pytestmark = pytest.mark.gpu
REC_RTOL = 1e-6


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float32))).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def case(oracle_mod):
    p = synth.workload_params("small")   # nt_interval = 2
    vt = synth.gradient_anomaly_model(p.nz, p.nx, seed=12)
    vs = synth.smooth(vt, 3.0)
    obs = oracle_mod.forward(p, vt)
    syn_o, grad_o = oracle_mod.fwi(p, vs, obs, run_fwi=0)
    assert np.linalg.norm(grad_o) > 0
    return p, vt, vs, obs, syn_o, grad_o


def _check(tag, syn, grad, syn_o, grad_o):
    print(f"[{tag}] rec rel={rel_l2(syn, syn_o):.3e} exact={np.mean(syn == syn_o):.4f} "
          f"grad rel={rel_l2(grad, grad_o):.3e}", flush=True)
    assert rel_l2(syn, syn_o) <= REC_RTOL
    assert rel_l2(grad, grad_o) <= GRAD_RTOL


def test_plain_run_fwi_0(case):
    p, vt, vs, obs, syn_o, grad_o = case
    syn, grad = solver.fwi_plain(p, dev(vs.reshape(-1)), dev(obs), run_fwi=0)
    _check("plain", host(syn), host(grad), syn_o, grad_o)


@pytest.mark.parametrize("ckpt", [0, 40])
def test_batch_run_fwi_0(case, ckpt):
    p, vt, vs, obs, syn_o, grad_o = case
    vel = dev(np.stack([vs.reshape(-1), vt.reshape(-1)]))
    syn, grad, mis = (host(t) for t in solver.fwi_batch(p, vel, dev(obs), run_fwi=0, ckpt_interval=ckpt))
    _check(f"batch ckpt={ckpt}", syn[0], grad[0], syn_o, grad_o)
    # the misfit is still 0.5*sum((syn-obs)^2)
    assert rel_err(mis[0], 0.5 * np.sum((syn_o.astype(np.float64) - obs) ** 2)) <= MISFIT_RTOL


def test_autograd_run_fwi_0(case):
    p, vt, vs, obs, syn_o, grad_o = case
    vel = dev(vs.reshape(1, -1)).requires_grad_()
    mis = solver.fwi2d_misfit(vel, dev(obs), p, ckpt_interval=40, run_fwi=0)
    mis.sum().backward()
    assert rel_l2(host(vel.grad)[0], -grad_o) <= GRAD_RTOL


def test_aco2d_and_host_entry_point_is_fwi_0(case, tmp_path):
    p, vt, vs, obs, syn_o, grad_o = case
    path = str(tmp_path / "input_params.txt")
    p.write(path)
    syn, grad = aco2d.fwi(vs.reshape(-1).astype(np.float64), obs, paramfile=path, is_fwi=0)
    _check("aco2d", syn, grad, syn_o, grad_o)
    lib = _lib.load()
    fp = ctypes.POINTER(ctypes.c_float)
    vel = np.ascontiguousarray(vs.reshape(-1), np.float32)
    syn2 = np.zeros(p.record_len, np.float32)
    grad2 = np.zeros(p.ni, np.float32)
    o = np.ascontiguousarray(obs)
    lib.fwi_2D(path.encode(), vel.ctypes.data_as(fp), syn2.ctypes.data_as(fp), o.ctypes.data_as(fp),
               grad2.ctypes.data_as(fp), 0, 0)
    assert lib.gpvi_compat_last_status() == 0
    assert np.array_equal(syn2, syn) and np.array_equal(grad2, grad)


@pytest.mark.parametrize("run_fwi", [0, 1])
@pytest.mark.parametrize("k", [1, 2])
def test_point_geometry_reversed_residual(oracle_mod, run_fwi, k):
    """One source, one receiver, residual = time-reversed wavelet: the
    oracle's closed-form case, through the batched solver."""
    v = hetero_model()
    for s, r in ((A, B), (B, A)):
        p = point_params(s, r, 251, k)
        obs = reversed_residual_obs(oracle_mod, p, v, run_fwi, np.float32)
        syn_o, grad_o = oracle_mod.fwi(p, v, obs, run_fwi=run_fwi)
        syn, grad, _ = (host(t) for t in solver.fwi_batch(p, dev(v.reshape(1, -1)), dev(obs), run_fwi=run_fwi))
        _check(f"point {s}->{r} run_fwi={run_fwi} k={k}", syn[0], grad[0], syn_o, grad_o)


def test_adjoint_keeps_the_forward_layout(case, oracle_mod):
    """forward() under one launch layout, backward() after the tuning and
    the time blocking changed: the gradient is the oracle's, because the
    adjoint phase uses the layout its forward phase stamped."""
    p, vt, vs, obs, _, _ = case
    _, grad_ref = oracle_mod.fwi(p, vs, obs)
    saved_tuning, saved_tb = _lib.get_tuning(), _lib.get_time_blocking()
    try:
        for (v0, tb0), (v1, tb1) in (((4, 1), (3, 2)), ((3, 2), (2, 1)), ((1, 1), (5, 1))):
            _lib.set_tuning(v0, 0, saved_tuning[2])
            _lib.set_time_blocking(tb0)
            vel = dev(vs.reshape(1, -1)).requires_grad_()
            mis = solver.fwi2d_misfit(vel, dev(obs), p, ckpt_interval=40)
            _lib.set_tuning(v1, 0, saved_tuning[2])
            _lib.set_time_blocking(tb1)
            mis.sum().backward()
            err = rel_l2(host(vel.grad)[0], -grad_ref)
            print(f"layout ({v0}, tb {tb0}) -> ({v1}, tb {tb1}): grad rel {err:.3e}", flush=True)
            assert err <= GRAD_RTOL
    finally:
        _lib.set_tuning(*saved_tuning)
        _lib.set_time_blocking(saved_tb)


def test_adjoint_without_forward_is_refused(case):
    p, vt, vs, obs, _, _ = case
    ws = torch.empty(solver.batch_workspace_bytes(p, 1, 0), dtype=torch.uint8, device="cuda")
    syn = torch.zeros(1, p.record_len, device="cuda")
    grad = torch.zeros(1, p.ni, device="cuda")
    with pytest.raises(_lib.GpviError) as exc:
        solver._batch_call(p, 1, None, dev(obs), syn, grad, None, 1, 0, _lib.GPVI_PHASE_ADJOINT, 0, ws, ws.device)
    assert exc.value.code == -1
