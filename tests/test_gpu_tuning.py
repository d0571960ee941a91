"""Launch shapes never change results.

Every tile variant, grid and steps-per-launch setting only changes which
wavefront computes which points; records, gradients and misfits must be
bit-identical across all of them and match the CPU oracle to the parity
bars.  Also runs the autotuner on a small case.
"""
import numpy as np
import pytest
import torch

from helpers import GRAD_RTOL, MISFIT_RTOL, case, rel_err, rel_l2
from geopvi_amd import _lib, autotune
from geopvi_amd.forward.fwi2d import solver

pytestmark = pytest.mark.gpu
REC_RTOL = 1e-6


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float32))).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def small(oracle_mod):
    # odd nt: a two-step launch schedule ends every pass with a one-step macro;
    # 100 columns: one-step launches use side-band items, two-step launches
    # do not, and both must give the same bits
    p, vt, vs = case("small", seed=11, nt=301, nt_interval=2, nx=100, nr=50)
    obs = oracle_mod.forward(p, vt)
    syn_o, grad_o = oracle_mod.fwi(p, vs, obs)
    vel = np.stack([vs.reshape(-1), (0.5 * (vs + vt)).reshape(-1)])
    return p, vel, obs, syn_o, grad_o


@pytest.fixture()
def restore_tuning():
    saved = (_lib.get_tuning(), _lib.get_time_blocking(), _lib.get_lanes())
    yield
    (v, g, s), tb, lanes = saved
    _lib.set_tuning(v, g, s)
    _lib.set_time_blocking(tb)
    _lib.set_lanes(lanes)


@pytest.mark.parametrize("ckpt", [0, 60])
def test_all_launch_shapes_bit_identical(small, oracle_mod, restore_tuning, ckpt):
    p, vel, obs, syn_o, grad_o = small
    vel_t, obs_t = dev(vel), dev(obs)
    ref = None
    for variant in range(-1, 6):   # -1: chosen per call
        for grid in (0, 8, 64):
            for tb in (1, 2):
                for graph_steps in ((32,) if grid else (32, 5)):
                    _lib.set_tuning(variant, grid, graph_steps)
                    _lib.set_time_blocking(tb)
                    out = [host(t) for t in solver.fwi_batch(p, vel_t, obs_t, ckpt_interval=ckpt)]
                    tag = f"variant={variant} grid={grid} tb={tb} seg={graph_steps} ckpt={ckpt}"
                    print(tag, flush=True)
                    if ref is None:
                        ref = out
                        syn, grad, mis = out
                        assert rel_l2(syn[0], syn_o) <= REC_RTOL
                        assert rel_l2(grad[0], grad_o) <= GRAD_RTOL
                        assert rel_err(mis[0], oracle_mod.misfit_f32(syn_o, obs)) <= MISFIT_RTOL
                    else:
                        for a, b in zip(out, ref):
                            assert np.array_equal(a, b), tag


def test_autotune_small(small, restore_tuning):
    p, vel, obs, _, _ = small
    cands = autotune.candidates(variants=(0, 4), grids=(0, 64), steps_per_launch=(1, 2))
    res = autotune.autotune(p, 2, cands=cands, steps=40, cache_path=None)
    assert (res["variant"], res["grid"], res["tb"]) in cands
    assert res["us_per_step"] > 0 and not res["cached"]
    assert len(res["timings"]) == len(cands)
    # applied process-wide
    v, g, _ = _lib.get_tuning()
    assert (v, g, _lib.get_time_blocking()) == (res["variant"], res["grid"], res["tb"])


@pytest.mark.parametrize("ckpt", [0, 60])
def test_lanes_bit_identical(small, restore_tuning, ckpt):
    """Units advanced on 1 to 4 streams (groups of samples, or the two shot
    halves of one sample), with graph segments or one launch per step, give
    the bits of a single lane."""
    p, vel, obs, _, _ = small
    assert p.ns == 3   # odd: the shot halves of one sample are uneven
    vel3 = np.concatenate([vel, vel[:1] * 1.01])
    obs = solver.forward_batch(p, dev(vel[1:2]))[0].contiguous()
    for nsample in (1, 2, 3):
        vel_t = dev(vel3[:nsample])
        ref = None
        for lanes in (1, 2, 3, 4):
            for graph in (True, False):
                _lib.set_lanes(lanes)
                out = [host(t) for t in solver.fwi_batch(p, vel_t, obs, ckpt_interval=ckpt, graph=graph)]
                tag = f"nsample={nsample} lanes={lanes} graph={graph} ckpt={ckpt}"
                print(tag, flush=True)
                if ref is None:
                    ref = out
                else:
                    for a, b in zip(out, ref):
                        assert np.array_equal(a, b), tag


def test_lanes_phase_split_and_one_stream(small, restore_tuning):
    """The autograd path (forward and adjoint phases in separate calls) and
    GPVI_FLAG_ONE_STREAM give the single-lane bits."""
    p, vel, obs, _, _ = small
    vel_t, obs_t = dev(vel), dev(obs)
    _lib.set_lanes(1)
    syn1, grad1, mis1 = (host(t) for t in solver.fwi_batch(p, vel_t, obs_t, ckpt_interval=60))
    _lib.set_lanes(2)
    v = vel_t.clone().requires_grad_()
    loss = solver.fwi2d_misfit(v, obs_t, p, ckpt_interval=60)
    loss.sum().backward()
    assert np.array_equal(host(loss), mis1.astype(host(loss).dtype))
    assert np.array_equal(host(v.grad), -grad1)
    n = vel.shape[0]
    syn = torch.empty(n, p.record_len, device="cuda")
    grad = torch.zeros(n, p.ni, device="cuda")
    mis = torch.empty(n, dtype=torch.float64, device="cuda")
    ws = torch.empty(solver.batch_workspace_bytes(p, n, 60), dtype=torch.uint8, device="cuda")
    solver._batch_call(p, n, vel_t, obs_t, syn, grad, mis, 1, 60, _lib.GPVI_PHASE_ALL, _lib.GPVI_FLAG_ONE_STREAM,
                       ws, vel_t.device)
    assert np.array_equal(host(syn), syn1) and np.array_equal(host(grad), grad1)
    assert np.array_equal(host(mis), mis1)


def test_fallback_division_same_bits(small, monkeypatch):
    """GPVI_NO_FASTDIV sends every wave through the fallback division (the
    double product RN_f(RN_d(a * RN_d(1/b)))) instead of the reciprocal
    sequence: the same bits, forward and adjoint, full store and checkpoints."""
    p, vel, obs, _, _ = small
    vel_t, obs_t = dev(vel), dev(obs)
    for ckpt in (0, 60):
        fast = [host(t) for t in solver.fwi_batch(p, vel_t, obs_t, ckpt_interval=ckpt)]
        monkeypatch.setenv("GPVI_NO_FASTDIV", "1")
        slow = [host(t) for t in solver.fwi_batch(p, vel_t, obs_t, ckpt_interval=ckpt)]
        monkeypatch.delenv("GPVI_NO_FASTDIV")
        for a, b in zip(fast, slow):
            assert np.array_equal(a, b), ckpt
