"""The reference's own example problem on the GPU (examples/fwi2d/fwi2d.py:
106-168 with examples/fwi2d/input/config.ini): 110 x 250 grid at 20 m,
12 shots at depth 1, 250 receivers at depth 9, 2001 steps of 2 ms, a
10-row water layer fixed at 1950 m/s, prior bounds from prior.txt.

Models come from the saved ADVI posterior the reference ships
(Linear_diagonal_parameter.npy through Real2Constr, see
tests/test_reference_inputs.py): its mean as the starting model and draws
z = u + exp(diag) * x as samples.  The observed data (the example's
waveform.npy is not in the reference) are modelled from the mean plus a
smooth seeded perturbation.

* The shots-in-turn solver, aco2d.fwi (a one-sample batch) and the batched
  solver on the first and last shot against the oracle: records bit-exact,
  misfit 1e-5, gradient 1e-4.
* One sample over all 12 shots through the batched solver against the
  oracle, and Posterior.log_prob with the example's mask on the device
  against its host path.
"""
import configparser

import numpy as np
import pytest
import torch

from helpers import GRAD_RTOL, MISFIT_RTOL, rel_err, rel_l2
from geopvi_amd import io, synth
from geopvi_amd.forward.fwi2d import aco2d, solver
from geopvi_amd.forward.fwi2d.posterior import Posterior
from geopvi_amd.nfvi import Real2Constr
from geopvi_amd.prior import Uniform
from test_reference_inputs import INPUTS, example_config, example_model

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float32))).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def example():
    p, vmean, mask, sigma = example_model()
    cfg = example_config()
    layer = cfg.getint("FWI", "layer_fixed")
    lo, up = io.load_prior_bounds(f"{INPUTS}/prior.txt", p.nz, p.nx, layer)
    rng = np.random.default_rng(17)
    pert = synth.smooth(rng.standard_normal((p.nz, p.nx)), 6.0)
    truth = vmean.copy()
    truth[layer:] = np.clip(vmean[layer:] + 150.0 * pert[layer:] / np.abs(pert).max(),
                            lo.reshape(-1, p.nx) + 1, up.reshape(-1, p.nx) - 1)
    obs = host(solver.forward_batch(p, dev(truth.reshape(1, -1)))[0])
    # four draws of the saved posterior
    param = np.load(f"{INPUTS}/Linear_diagonal_parameter.npy", allow_pickle=False)
    u = param[:25000]
    x = np.random.default_rng(18).standard_normal((4, 25000))
    with torch.no_grad():
        z, _ = Real2Constr(25000, lower=lo, upper=up).forward(torch.from_numpy(u + sigma * x))
    samples = np.repeat(vmean.reshape(1, -1), 4, axis=0)
    samples[:, mask.reshape(-1)] = z.numpy()
    print(f"example: {p.nz}x{p.nx}, {p.ns} shots, {p.nt} steps, dx {p.dx}, dt {p.dt}", flush=True)
    return p, vmean, mask, samples, obs, lo, up


def _two_shots(p, obs):
    q = p.replace(ns=2, ds=(p.ns - 1) * p.ds)   # shots 0 and 11
    o = np.ascontiguousarray(obs.reshape(p.ns, p.nt, p.nr)[[0, -1]]).reshape(-1)
    return q, o


def test_first_and_last_shot_plain_and_batched(example, oracle_mod, tmp_path):
    p, vmean, mask, samples, obs, _, _ = example
    q, oq = _two_shots(p, obs)
    path = str(tmp_path / "input_params.txt")
    q.write(path)
    syn_o, grad_o = oracle_mod.fwi(q, samples[0], oq)
    m_o = oracle_mod.misfit_f32(syn_o, oq)
    syn_p, grad_p = (host(t) for t in solver.fwi_plain(q, dev(samples[0]), dev(oq)))
    print(f"shots in turn: grad rel {rel_l2(grad_p, grad_o):.2e}", flush=True)
    assert np.array_equal(syn_p, syn_o) and rel_l2(grad_p, grad_o) <= GRAD_RTOL
    syn, grad = aco2d.fwi(samples[0].astype(np.float64), oq, paramfile=path)
    print(f"aco2d.fwi: grad rel {rel_l2(grad, grad_o):.2e}", flush=True)
    assert np.array_equal(syn, syn_o) and rel_l2(grad, grad_o) <= GRAD_RTOL
    syn_b, grad_b, mis_b = (host(t) for t in solver.fwi_batch(q, dev(samples), dev(oq)))
    print(f"batched: grad rel {rel_l2(grad_b[0], grad_o):.2e}, misfit rel {rel_err(mis_b[0], m_o):.2e}", flush=True)
    assert np.array_equal(syn_b[0], syn_o)
    assert rel_err(mis_b[0], m_o) <= MISFIT_RTOL and rel_l2(grad_b[0], grad_o) <= GRAD_RTOL
    syn_o3, grad_o3 = oracle_mod.fwi(q, samples[3], oq)
    assert np.array_equal(syn_b[3], syn_o3) and rel_l2(grad_b[3], grad_o3) <= GRAD_RTOL


def test_all_shots_and_posterior_with_the_water_mask(example, oracle_mod, tmp_path):
    p, vmean, mask, samples, obs, lo, up = example
    syn_o, grad_o = oracle_mod.fwi(p, samples[1], obs)
    syn_b, grad_b, mis_b = (host(t) for t in solver.fwi_batch(p, dev(samples[1:2]), dev(obs)))
    print(f"12 shots: grad rel {rel_l2(grad_b[0], grad_o):.2e}", flush=True)
    assert np.array_equal(syn_b[0], syn_o)
    assert rel_err(mis_b[0], oracle_mod.misfit_f32(syn_o, obs)) <= MISFIT_RTOL
    assert rel_l2(grad_b[0], grad_o) <= GRAD_RTOL

    cfg = example_config()
    prior = Uniform(lower=lo, upper=up)
    post = Posterior(obs, cfg, vel_fixed=cfg.getfloat("FWI", "vel_fixed"), sigma=0.1, log_prior=prior.log_prob,
                     mask=mask.reshape(-1), paramfile=str(tmp_path / "input_params.txt"))
    x = samples[:2][:, mask.reshape(-1)]
    x_cpu = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    lp_cpu = post.log_prob(x_cpu)
    lp_cpu.sum().backward()
    x_dev = torch.tensor(x, dtype=torch.float64, device="cuda", requires_grad=True)
    lp_dev = post.log_prob(x_dev)
    lp_dev.sum().backward()
    np.testing.assert_allclose(lp_dev.detach().cpu().numpy(), lp_cpu.detach().numpy(), rtol=1e-9)
    assert rel_l2(x_dev.grad.cpu().numpy(), x_cpu.grad.numpy()) <= 1e-5
    # the host path's likelihood of sample 1 is the oracle's
    loss1 = -(lp_cpu[1].item() - prior.log_prob(x_cpu[1:2]).item()) * 0.1 ** 2
    assert rel_err(loss1, oracle_mod.misfit_f32(syn_o, obs)) <= MISFIT_RTOL
