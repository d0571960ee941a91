"""The end-to-end inversion case at its full size (shared by the fixture
script make_c4_fixtures.py and tests/test_gpu_c4.py).

ADVI (diagonal ``Linear``) and PSVI (``Linear`` with the 12 bands of a 5x5
correlation template) on the 200x400 grid (80 000 parameters), 20 shots,
2000 steps, a ``Real2Constr`` bound of [1400, 4800] m/s and a ``Uniform``
prior over the same box; 32 samples per iteration, Adam with lr 0.002
(examples/fwi2d/fwi2d.py:61, geopvi/nfvi/models.py:113-122,
geopvi/nfvi/flows.py:205-220).  Observed data: the true model's records
plus Gaussian noise at 1 % of their RMS; sigma = 1 % of the clean RMS.
Base draws come from one seeded CPU ``torch.Generator`` in float64, so the
CPU run and the device run see the same samples.
"""
from __future__ import annotations

import hashlib

import numpy as np

SEED = 1
NSAMPLE = 32
NITER = 2
LR = 0.002
LOWER, UPPER = 1400.0, 4800.0
METHODS = ("advi", "psvi")
NSUBSET = 8192          # flow-parameter and gradient entries kept in the fixture


def params():
    from geopvi_amd import synth
    return synth.workload_params("c1")


def models():
    """(true model (nz, nx) float64, starting mean (nz*nx,) float64)."""
    from geopvi_amd import synth
    vt = synth.true_model("c1", seed=SEED)
    vs = np.clip(synth.smooth(vt, 8.0).reshape(-1), LOWER + 50, UPPER - 50)
    return vt, vs


def observed(clean: np.ndarray):
    """(noisy observed records float32, sigma)."""
    from geopvi_amd import synth
    clean = np.asarray(clean, np.float32)
    sigma = 0.01 * float(np.sqrt(np.mean(clean.astype(np.float64) ** 2)))
    return synth.add_noise(clean, 0.01, seed=SEED + 200), sigma


def records_hash(rec) -> str:
    return hashlib.sha256(np.ascontiguousarray(rec, np.float32).tobytes()).hexdigest()


def flows(method: str, p, vs):
    """Fresh flows of one method, starting at the mean model ``vs``."""
    from geopvi_amd.nfvi import Linear, Real2Constr, psvi_mask_2D
    dim = p.ni
    u0 = np.log(vs - LOWER) - np.log(UPPER - vs)
    if method == "advi":
        lin = Linear(dim, kernel="diagonal", param=np.concatenate([u0, np.full(dim, 0.01)]))
    else:
        mask = psvi_mask_2D(np.ones((5, 5)), dim, nx=p.nx, nz=p.nz)
        lin = Linear(dim, kernel="structured", mask=mask,
                     param=np.concatenate([u0, np.full(dim, 0.01), np.zeros(mask.size)]))
    return [lin, Real2Constr(dim, lower=LOWER, upper=UPPER)]


def prior(p):
    from geopvi_amd.prior import Uniform
    return Uniform(lower=np.full(p.ni, LOWER), upper=np.full(p.ni, UPPER))


def generator(method: str):
    import torch
    return torch.Generator().manual_seed(SEED + (0 if method == "advi" else 1))


def param_vector(q) -> np.ndarray:
    """Every flow parameter of ``q`` in named_parameters order, float64."""
    return np.concatenate([t.detach().cpu().numpy().reshape(-1).astype(np.float64)
                           for _, t in q.named_parameters()])


def subset(n: int, salt: int) -> np.ndarray:
    """Sorted, seeded indices of NSUBSET entries of a vector of length n."""
    rng = np.random.default_rng(1000 + salt)
    return np.sort(rng.choice(n, size=min(NSUBSET, n), replace=False))
