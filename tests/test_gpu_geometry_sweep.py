"""Seeded random geometries against the CPU oracle.

Each case draws a grid size, absorbing width, stencil half-width (Lc 3..6),
store interval (1..3), checkpoint interval, shot and receiver spreads and
depths anywhere in the padded grid (absorbing bands and edge rows included).
The padded width decides the side-band widths and the row shift of the
field layout, so the cases cover many combinations of strips, bands and
alignment.  Records must equal the oracle's bit for bit (batched, with and
without checkpoints, and plain); misfit and gradient meet the parity bars.
"""
import numpy as np
import pytest
import torch

from helpers import GRAD_RTOL, MISFIT_RTOL, case, rel_err, rel_l2
from geopvi_amd.forward.fwi2d import solver

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1))).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def draw(seed):
    rng = np.random.default_rng(1000 + seed)
    nx, nz = int(rng.integers(24, 161)), int(rng.integers(16, 81))
    lc, pml0, k = int(rng.integers(3, 7)), int(rng.integers(2, 21)), int(rng.integers(1, 4))
    pml = pml0 + lc + 1
    ntx, ntz = nx + 2 * pml, nz + 2 * pml
    ns = int(rng.integers(1, 4))
    ds = int(rng.integers(1, max(2, (ntx - 1) // max(ns - 1, 1)))) if ns > 1 else 1
    ns0 = int(rng.integers(-pml, ntx - pml - (ns - 1) * ds))
    nr = int(rng.integers(1, 40))
    dr = int(rng.integers(1, max(2, (ntx - 1) // max(nr - 1, 1)))) if nr > 1 else 1
    nr0 = int(rng.integers(-pml, ntx - pml - (nr - 1) * dr))
    over = dict(nx=nx, nz=nz, Lc=lc, pml0=pml0, nt_interval=k, nt=int(rng.integers(120, 261)), ns=ns, ds=ds,
                ns0=ns0, nr=nr, dr=dr, nr0=nr0, depths=int(rng.integers(-pml, ntz - pml)),
                depthr=int(rng.integers(-pml, ntz - pml)))
    ckpt = int(rng.integers(10, 80))
    return over, ckpt


def misfit_ok(mis, syn_o, obs, oracle_mod):
    """The device misfit (float differences, squared and summed in double)
    against the same sum over the oracle's records, and against the
    reference's float32 sum: equal within the bar, or both below float32's
    range when no wave reached a receiver (the float32 squares underflow to
    0)."""
    d = (np.asarray(syn_o, np.float32) - np.asarray(obs, np.float32)).astype(np.float64)
    mis64 = 0.5 * float(np.sum(d * d))
    ok64 = abs(mis - mis64) <= 1e-10 * mis64 if mis64 > 0 else mis == 0.0
    ref = oracle_mod.misfit_f32(syn_o, obs)
    ok32 = rel_err(mis, ref) <= MISFIT_RTOL if ref != 0 else np.float32(mis) == 0
    return ok64 and ok32


@pytest.mark.parametrize("seed", range(24))
def test_random_geometry_matches_oracle(oracle_mod, seed):
    over, ckpt = draw(seed)
    p, vt, vs = case("small", seed=seed, **over)
    print(f"sweep {seed}: {over} ckpt={ckpt} padded {p.ntz}x{p.ntx}", flush=True)
    obs = oracle_mod.forward(p, vt)
    syn_o, grad_o = oracle_mod.fwi(p, vs, obs)
    vel = dev(np.stack([vs.reshape(-1), vt.reshape(-1)])).view(2, -1)
    for ck in (0, ckpt):
        syn_b, grad_b, mis_b = (host(t) for t in solver.fwi_batch(p, vel, dev(obs), ckpt_interval=ck))
        assert np.array_equal(syn_b[0], syn_o), f"records, ckpt={ck}"
        assert rel_l2(grad_b[0], grad_o) <= GRAD_RTOL
        assert misfit_ok(float(mis_b[0]), syn_o, obs, oracle_mod)
    syn_p, grad_p = (host(t) for t in solver.fwi_plain(p, dev(vs), dev(obs)))
    assert np.array_equal(syn_p, syn_o)
    assert rel_l2(grad_p, grad_o) <= GRAD_RTOL
