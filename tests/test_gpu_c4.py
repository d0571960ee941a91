"""End-to-end variational inversion at its full size on the device, against
a CPU run of the same case with the C oracle as likelihood
(tests/golden/c4_case.py; fixture tests/golden/c4_inversion.npz written by
tests/golden/make_c4_fixtures.py).

ADVI and PSVI on the 200x400 grid (80 000 parameters), 20 shots x 2000
steps, 32 samples per iteration, 2 Adam iterations from the same seeded
base draws.  The device run keeps flows, samples and gradients on the GPU
and takes the likelihood from the batched HIP solver (checkpointed adjoint)
through Posterior's device path.

* The observed data modelled on the GPU hash to the oracle's records (the
  full 20-shot, 2000-step forward bit for bit).
* Per iteration: every sample's misfit within 1e-5 of the oracle's; the
  -ELBO within 1e-5; d(-ELBO)/dz averaged over the samples within 1e-4
  (relative L2 over the fixture's cells and in norm) at the first
  iteration, where both runs evaluate identical samples.  The second
  iteration's samples come from parameters that already went through one
  Adam step, which rounds differently where a gradient component is within
  rounding of zero (see below); its gradient is held to 1e-3.
* The flow parameters after the two steps.  Adam moves each parameter by
  lr * m/sqrt(v): where a gradient component is large, both runs move it
  by the same amount to ~1e-9 lr; where it is at the rounding level of its
  own sum over samples and cells (PSVI's band entries in particular), the
  two runs' ratios m/sqrt(v) differ by up to a few percent of lr (0.04 lr
  measured; 98.8 % of PSVI's entries and 99.9 % or more of ADVI's within
  1e-3 lr).
  So: at least 95 % of the fixture's entries within 1e-3 lr, all within
  0.25 lr (a component moved the other way would be 2 lr apart), and the
  norm of the change within 1e-5.
"""
import copy
import os

import numpy as np
import pytest
import torch

from helpers import fwi_config, rel_l2
from geopvi_amd.engine import best_ckpt_interval
from geopvi_amd.forward.fwi2d import solver
from geopvi_amd.forward.fwi2d.posterior import Posterior
from geopvi_amd.nfvi import FlowsBasedDistribution, VariationalInversion
from golden import c4_case as C

pytestmark = pytest.mark.gpu
FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c4_inversion.npz")
DEV = torch.device("cuda", 0)


@pytest.fixture()
def float64_default():
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


@pytest.fixture(scope="module")
def problem(tmp_path_factory):
    ref = dict(np.load(FIXTURE, allow_pickle=False))
    p = C.params()
    vt, vs = C.models()
    clean = solver.forward_batch(p, torch.from_numpy(vt.astype(np.float32).reshape(1, -1)).to(DEV))[0]
    clean = clean.cpu().numpy()
    obs, sigma = C.observed(clean)
    path = str(tmp_path_factory.mktemp("c4") / "input_params.txt")
    return p, vs, clean, obs, sigma, path, ref


def test_observed_data_are_the_oracles(problem):
    p, vs, clean, obs, sigma, path, ref = problem
    assert C.records_hash(clean) == str(ref["obs_sha256"])
    assert sigma == float(ref["sigma"])


@pytest.mark.parametrize("salt,method", list(enumerate(C.METHODS)))
def test_inversion_matches_the_cpu_oracle_run(problem, float64_default, salt, method):
    p, vs, clean, obs, sigma, path, ref = problem
    prior = C.prior(p)
    post = Posterior(obs, fwi_config(p), sigma=sigma, log_prior=prior.log_prob, paramfile=path,
                     ckpt_interval=best_ckpt_interval(p))
    sub_z = C.subset(p.ni, salt)
    rec = {"misfit": [], "zgrad_sub": [], "zgrad_norm": []}

    def log_post(z):
        def hook(g):
            gm = g.mean(dim=0).cpu().numpy()
            rec["zgrad_sub"].append(gm[sub_z])
            rec["zgrad_norm"].append(float(np.linalg.norm(gm)))
        z.register_hook(hook)
        loss = post._misfit_device(z)
        rec["misfit"].append(loss.detach().cpu().numpy())
        return -loss / sigma ** 2 + prior.log_prob(z)

    q = FlowsBasedDistribution(copy.deepcopy(C.flows(method, p, vs)), generator=C.generator(method)).to(DEV)
    theta0 = C.param_vector(q)
    loss = VariationalInversion(q, log_post).update(lr=C.LR, n_iter=C.NITER, nsample=C.NSAMPLE)
    d = C.param_vector(q) - theta0
    sub_t = C.subset(d.size, 10 + salt)

    mis = np.stack(rec["misfit"])
    m_rel = np.abs(mis - ref[f"{method}_misfit"]) / np.abs(ref[f"{method}_misfit"])
    l_rel = np.abs(np.array(loss) - ref[f"{method}_loss"]) / np.abs(ref[f"{method}_loss"])
    g_rel = [rel_l2(a, b) for a, b in zip(rec["zgrad_sub"], ref[f"{method}_zgrad_sub"])]
    n_rel = np.abs(np.array(rec["zgrad_norm"]) - ref[f"{method}_zgrad_norm"]) / ref[f"{method}_zgrad_norm"]
    dd = np.abs(d[sub_t] - ref[f"{method}_dtheta_sub"])
    frac = float(np.mean(dd <= 1e-3 * C.LR))
    dn_rel = abs(np.linalg.norm(d) - float(ref[f"{method}_dtheta_norm"])) / float(ref[f"{method}_dtheta_norm"])
    print(f"{method}: misfit rel max {m_rel.max():.2e}, -ELBO rel {l_rel.max():.2e}, dz gradient rel {g_rel} "
          f"(norm {n_rel.max():.2e}), parameters within 1e-3 lr: {frac:.5f}, max diff {dd.max() / C.LR:.3f} lr, "
          f"change norm rel {dn_rel:.2e}", flush=True)
    assert d.size == int(ref[f"{method}_ntheta"])
    assert m_rel.max() <= 1e-5 and l_rel.max() <= 1e-5
    assert g_rel[0] <= 1e-4 and n_rel[0] <= 1e-4
    assert g_rel[1] <= 1e-3 and n_rel[1] <= 1e-3
    assert frac >= 0.95 and dd.max() <= 0.25 * C.LR and dn_rel <= 1e-5
