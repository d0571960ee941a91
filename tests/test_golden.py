"""Committed golden fixtures (tests/golden/*.npz, written by make_golden.py).

CPU: the oracle still reproduces them bit for bit.
GPU: the HIP plain and batched paths match them -- records bit-exact, misfit
within 1e-5 relative, gradient within 1e-4 relative L2 (also stated against
the double twin).
"""
import glob
import os

import numpy as np
import pytest

from helpers import GRAD_RTOL, MISFIT_RTOL, rel_err, rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
# the solver fixtures of make_golden.py (c4_inversion.npz is the inversion
# fixture of test_gpu_c4.py, with its own layout)
FIXTURES = sorted(f for f in glob.glob(os.path.join(HERE, "golden", "*.npz")) if not os.path.basename(f).startswith("c4_"))


def load(path):
    from golden.make_golden import params_from_vector
    d = dict(np.load(path, allow_pickle=False))
    return params_from_vector(d["params"]), d


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(f)[:-4] for f in FIXTURES])
def test_oracle_reproduces_fixture(oracle_mod, path):
    p, d = load(path)
    obs = oracle_mod.forward(p, d["vel_true"])
    assert np.array_equal(obs, d["obs"])
    syn, grad = oracle_mod.fwi(p, d["vel"], d["obs"])
    assert np.array_equal(syn, d["syn"])
    assert np.array_equal(grad, d["grad"])
    assert oracle_mod.misfit_f32(syn, d["obs"]) == d["misfit"]


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(f)[:-4] for f in FIXTURES])
def test_hip_matches_fixture(path):
    import torch
    from geopvi_amd.forward.fwi2d import solver
    p, d = load(path)
    dev = torch.device("cuda", 0)
    vel = torch.from_numpy(d["vel"].reshape(-1)).to(dev)
    obs = torch.from_numpy(d["obs"]).to(dev)
    fwd = solver.forward_plain(p, torch.from_numpy(d["vel_true"].reshape(-1)).to(dev))
    syn_p, grad_p = solver.fwi_plain(p, vel, obs)
    syn_b, grad_b, mis_b = solver.fwi_batch(p, vel.view(1, -1), obs)
    torch.cuda.synchronize()
    fwd, syn_p, grad_p = fwd.cpu().numpy(), syn_p.cpu().numpy(), grad_p.cpu().numpy()
    syn_b, grad_b, mis_b = syn_b.cpu().numpy()[0], grad_b.cpu().numpy()[0], float(mis_b.cpu()[0])
    print(f"[{os.path.basename(path)}] plain grad rel {rel_l2(grad_p, d['grad']):.2e} "
          f"(vs f64 {rel_l2(grad_p, d['grad64']):.2e}); batch grad rel {rel_l2(grad_b, d['grad']):.2e}; "
          f"misfit rel {rel_err(mis_b, d['misfit']):.2e}")
    assert np.array_equal(fwd, d["obs"])
    assert np.array_equal(syn_p, d["syn"])
    assert np.array_equal(syn_b, d["syn"])
    assert rel_l2(grad_p, d["grad"]) <= GRAD_RTOL
    assert rel_l2(grad_b, d["grad"]) <= GRAD_RTOL
    assert rel_l2(grad_b, d["grad64"]) <= GRAD_RTOL
    assert rel_err(mis_b, d["misfit"]) <= MISFIT_RTOL
