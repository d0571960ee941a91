"""The headline workload at full size: 16 samples x 20 shots on the
200 x 400 grid, 200 receivers, 2000 steps; observed data with additive
Gaussian noise at 1 % of their RMS.

* Checkpointed (every 100 steps) and full-store batches give the same bits
  for all 16 samples.
* Records of the first and last shot of sample 0 are bit-identical to the
  CPU oracle's.
* On the two-shot subset of the same geometry (shots 0 and 19), misfit and
  gradient match the oracle to the parity bars, and the two-shot gradient
  is the sum of the one-shot gradients (shot additivity).
"""
import numpy as np
import pytest
import torch

from helpers import GRAD_RTOL, MISFIT_RTOL, rel_err, rel_l2
from geopvi_amd import synth
from geopvi_amd.forward.fwi2d import solver

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float32))).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def c1():
    p = synth.workload_params("c1")
    vt = synth.true_model("c1", seed=1)
    samples = synth.gaussian_samples(synth.smooth(vt, 8.0), 50.0, 16, seed=101).astype(np.float32)
    clean = host(solver.forward_batch(p, dev(vt).view(1, -1))[0])
    obs = dev(synth.add_noise(clean, 0.01, seed=201))   # 1 % RMS noise, as the workload specifies
    print(f"c1: {p.nz}x{p.nx}, {p.ns} shots, {p.nt} steps, 16 samples, 1 % noise", flush=True)
    return p, vt, samples, obs


def test_checkpointed_and_full_store_batches_identical(c1):
    p, _, samples, obs = c1
    vel = dev(samples)
    full = [host(t) for t in solver.fwi_batch(p, vel, obs, ckpt_interval=0)]
    ck = [host(t) for t in solver.fwi_batch(p, vel, obs, ckpt_interval=100)]
    for a, b in zip(full, ck):
        assert np.array_equal(a, b)
    assert np.all(np.isfinite(full[1])) and np.all(full[2] > 0)


def test_first_and_last_shot_against_oracle(c1, oracle_mod):
    p, _, samples, obs = c1
    v0 = samples[0]
    syn = host(solver.forward_batch(p, dev(v0).view(1, -1)))[0].reshape(p.ns, p.nt, p.nr)
    q = p.replace(ns=2, ds=(p.ns - 1) * p.ds)      # shots 0 and 19 of the geometry
    rec_o = oracle_mod.forward(q, v0).reshape(2, p.nt, p.nr)
    assert np.array_equal(syn[0], rec_o[0]) and np.array_equal(syn[-1], rec_o[1])

    obs_q = np.ascontiguousarray(host(obs).reshape(p.ns, p.nt, p.nr)[[0, -1]]).reshape(-1)
    syn_o, grad_o = oracle_mod.fwi(q, v0, obs_q)
    s2, g2, m2 = (host(t) for t in solver.fwi_batch(q, dev(v0).view(1, -1), dev(obs_q)))
    print(f"c1 shots 0,19: grad rel {rel_l2(g2[0], grad_o):.2e}, "
          f"misfit rel {rel_err(m2[0], oracle_mod.misfit_f32(syn_o, obs_q)):.2e}", flush=True)
    assert np.array_equal(s2[0], syn_o)
    assert rel_err(m2[0], oracle_mod.misfit_f32(syn_o, obs_q)) <= MISFIT_RTOL
    assert rel_l2(g2[0], grad_o) <= GRAD_RTOL

    # shot additivity: the two-shot gradient is the sum of the one-shot ones
    ga = host(solver.fwi_batch(q.replace(ns=1), dev(v0).view(1, -1), dev(obs_q[:p.nt * p.nr]))[1])[0]
    qb = q.replace(ns=1, ns0=q.ns0 + q.ds)
    gb = host(solver.fwi_batch(qb, dev(v0).view(1, -1), dev(obs_q[p.nt * p.nr:]))[1])[0]
    assert rel_l2(ga.astype(np.float64) + gb, g2[0]) <= 1e-6
