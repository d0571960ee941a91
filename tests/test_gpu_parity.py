"""HIP solver vs the CPU oracle on identical seeded inputs.

Bars (FP32, stated by the parity contract): misfit relative error <= 1e-5,
gradient relative L2 <= 1e-4.  Records are additionally held to 1e-6
relative L2 (same float operations in the same order as the oracle; the
only expected differences are the double-precision imaging coefficient and
the shot-parallel gradient sum of the batched path).
"""
import numpy as np
import pytest
import torch

from helpers import GRAD_RTOL, MISFIT_RTOL, case, rel_err, rel_l2
from geopvi_amd.forward.fwi2d import solver

pytestmark = pytest.mark.gpu
REC_RTOL = 1e-6


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1))).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


def _oracle_case(oracle_mod, name, seed=3, **kw):
    p, vt, vs = case(name, seed=seed, **kw)
    obs = oracle_mod.forward(p, vt)
    syn_o, grad_o = oracle_mod.fwi(p, vs, obs)
    return p, vt, vs, obs, syn_o, grad_o


@pytest.fixture(scope="module")
def tiny(oracle_mod):
    return _oracle_case(oracle_mod, "tiny")


def _report(tag, syn, syn_o, grad, grad_o):
    exact = float(np.mean(syn == syn_o))
    print(f"[{tag}] rec rel={rel_l2(syn, syn_o):.3e} exact={exact:.4f} grad rel={rel_l2(grad, grad_o):.3e}",
          flush=True)


def test_plain_forward_tiny(tiny, oracle_mod):
    p, vt, vs, obs, _, _ = tiny
    syn = host(solver.forward_plain(p, dev(vt)))
    _report("plain-forward", syn, obs, np.zeros(1), np.zeros(1))
    assert rel_l2(syn, obs) <= REC_RTOL


def test_plain_fwi_tiny(tiny, oracle_mod):
    p, vt, vs, obs, syn_o, grad_o = tiny
    syn, grad = solver.fwi_plain(p, dev(vs), dev(obs))
    syn, grad = host(syn), host(grad)
    _report("plain-fwi", syn, syn_o, grad, grad_o)
    assert rel_l2(syn, syn_o) <= REC_RTOL
    assert rel_err(oracle_mod.misfit_f32(syn, obs), oracle_mod.misfit_f32(syn_o, obs)) <= MISFIT_RTOL
    assert rel_l2(grad, grad_o) <= GRAD_RTOL


@pytest.mark.parametrize("ckpt", [0, 40, 64])
def test_batch_fwi_tiny(tiny, oracle_mod, ckpt):
    p, vt, vs, obs, syn_o, grad_o = tiny
    vel = dev(np.stack([vs.reshape(-1), vt.reshape(-1)]))
    syn, grad, mis = solver.fwi_batch(p, vel.view(2, -1), dev(obs), ckpt_interval=ckpt)
    syn, grad, mis = host(syn), host(grad), host(mis)
    _report(f"batch ckpt={ckpt}", syn[0], syn_o, grad[0], grad_o)
    assert rel_l2(syn[0], syn_o) <= REC_RTOL
    assert rel_err(mis[0], oracle_mod.misfit_f32(syn_o, obs)) <= MISFIT_RTOL
    assert rel_l2(grad[0], grad_o) <= GRAD_RTOL
    # second sample is the true model: zero residual up to rounding
    assert mis[1] <= 1e-6 * mis[0]


def test_batch_graph_matches_stepwise(tiny):
    p, vt, vs, obs, _, _ = tiny
    vel = dev(vs).view(1, -1)
    s1, g1, m1 = solver.fwi_batch(p, vel, dev(obs), graph=True)
    s2, g2, m2 = solver.fwi_batch(p, vel, dev(obs), graph=False)
    assert np.array_equal(host(s1), host(s2))
    assert np.array_equal(host(g1), host(g2))
    assert np.array_equal(host(m1), host(m2))


@pytest.mark.parametrize("lc", [3, 4, 5, 6])
def test_plain_fwi_orders(oracle_mod, lc):
    p, vt, vs, obs, syn_o, grad_o = _oracle_case(oracle_mod, "tiny", seed=5, Lc=lc, nt=200, nt_interval=1)
    syn, grad = solver.fwi_plain(p, dev(vs), dev(obs))
    syn, grad = host(syn), host(grad)
    _report(f"Lc={lc}", syn, syn_o, grad, grad_o)
    assert rel_l2(syn, syn_o) <= REC_RTOL
    assert rel_l2(grad, grad_o) <= GRAD_RTOL


def test_c0_plain_and_batch(oracle_mod):
    p, vt, _ = case("c0", seed=0)
    from geopvi_amd import synth
    vs = synth.smooth(vt, 8.0)
    obs = oracle_mod.forward(p, vt)
    syn_o, grad_o = oracle_mod.fwi(p, vs, obs)
    mis_o = oracle_mod.misfit_f32(syn_o, obs)
    syn, grad = solver.fwi_plain(p, dev(vs), dev(obs))
    syn, grad = host(syn), host(grad)
    _report("c0 plain", syn, syn_o, grad, grad_o)
    assert rel_err(oracle_mod.misfit_f32(syn, obs), mis_o) <= MISFIT_RTOL
    assert rel_l2(grad, grad_o) <= GRAD_RTOL
    sb, gb, mb = solver.fwi_batch(p, dev(vs).view(1, -1), dev(obs), ckpt_interval=100)
    sb, gb, mb = host(sb), host(gb), host(mb)
    _report("c0 batch ckpt=100", sb[0], syn_o, gb[0], grad_o)
    assert rel_err(mb[0], mis_o) <= MISFIT_RTOL
    assert rel_l2(gb[0], grad_o) <= GRAD_RTOL


@pytest.mark.parametrize("lc,k,ckpt", [(3, 2, 0), (3, 1, 40), (6, 2, 0), (4, 3, 50)])
def test_side_band_items(oracle_mod, lc, k, ckpt):
    """Grids wide enough for side-band work items (band width 16 for Lc 3-4,
    32 for Lc 6), with sources from the first to the last inner column and a
    receiver on every inner column, so injection and records also happen in
    band-item columns."""
    nx, ns = 100, 3
    p, vt, vs = case("small", seed=13, nx=nx, nz=30, Lc=lc, nt_interval=k, nt=240, ns=ns, ns0=0,
                     ds=(nx - 1) // (ns - 1), nr=nx, dr=1, nr0=0, depths=1, depthr=0)
    obs = oracle_mod.forward(p, vt)
    syn_o, grad_o = oracle_mod.fwi(p, vs, obs)
    syn_p, grad_p = solver.fwi_plain(p, dev(vs), dev(obs))
    vel = dev(np.stack([vs.reshape(-1), vt.reshape(-1)]))
    syn_b, grad_b, mis_b = solver.fwi_batch(p, vel.view(2, -1), dev(obs), ckpt_interval=ckpt)
    syn_p, grad_p, syn_b, grad_b, mis_b = host(syn_p), host(grad_p), host(syn_b), host(grad_b), host(mis_b)
    _report(f"bands Lc={lc} k={k} ckpt={ckpt}", syn_b[0], syn_o, grad_b[0], grad_o)
    assert np.array_equal(syn_p, syn_o)          # same float operations as the oracle
    assert np.array_equal(syn_b[0], syn_o)
    assert rel_l2(grad_p, grad_o) <= GRAD_RTOL
    assert rel_l2(grad_b[0], grad_o) <= GRAD_RTOL
    assert rel_err(mis_b[0], oracle_mod.misfit_f32(syn_o, obs)) <= MISFIT_RTOL


@pytest.mark.parametrize("ckpt", [0, 40])
def test_sources_and_receivers_inside_the_absorbing_bands(oracle_mod, ckpt):
    """The reference places sources and receivers anywhere in the padded
    grid (fdtd2D_modelling.c:125-152, no bounds checks): here the first
    source sits in the left band's copy columns (ix < Lc), the last one in
    the right band, all at a depth inside the top band, and the receivers
    span every padded column at a depth inside the bottom band (above its
    last Lc rows, which the reference never updates)."""
    p, vt, vs = case("small", seed=17, nx=100, nz=30, nt=520, ns=3)
    pml = p.pml0 + p.Lc + 1
    ntx, ntz = p.nx + 2 * pml, p.nz + 2 * pml
    s_first, s_last = 2, ntx - 3
    p = p.replace(ns0=s_first - pml, ds=(s_last - s_first) // (p.ns - 1), depths=6 - pml,
                  nr0=-pml, nr=ntx, dr=1, depthr=ntz - 7 - pml)
    obs = oracle_mod.forward(p, vt)
    syn_o, grad_o = oracle_mod.fwi(p, vs, obs)
    syn_p, grad_p = (host(t) for t in solver.fwi_plain(p, dev(vs), dev(obs)))
    syn_b, grad_b, mis_b = (host(t) for t in solver.fwi_batch(p, dev(vs).view(1, -1), dev(obs), ckpt_interval=ckpt))
    _report(f"in-band geometry ckpt={ckpt}", syn_b[0], syn_o, grad_b[0], grad_o)
    assert np.abs(syn_o).max() > 0
    assert np.array_equal(syn_p, syn_o) and np.array_equal(syn_b[0], syn_o)
    assert rel_l2(grad_p, grad_o) <= GRAD_RTOL and rel_l2(grad_b[0], grad_o) <= GRAD_RTOL
    assert rel_err(mis_b[0], oracle_mod.misfit_f32(syn_o, obs)) <= MISFIT_RTOL
