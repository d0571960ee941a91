"""Variational caller chain on the CPU: oracle pins and torch modules vs oracle."""
import numpy as np
import pytest
import torch

from oracle import nfvi_oracle as O
from geopvi_amd.nfvi import Constr2Real, FlowsBasedDistribution, Linear, Real2Constr, psvi_mask_2D
from geopvi_amd.prior import Uniform, smooth_matrix_2D

RNG = np.random.default_rng(5)


@pytest.fixture(autouse=True)
def float64_default():
    # the reference's inversion drivers run torch in float64
    # (examples/fwi2d/fwi2d.py:84)
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


# ---- pins of the oracle --------------------------------------------------

def test_psvi_mask_offsets_follow_template_geometry():
    nx, nz = 7, 5
    mask = psvi_mask_2D(np.ones((3, 3)), nx * nz, nx=nx, nz=nz)
    offs = (mask == 0).sum(axis=1)
    # entries before the centre of a 3x3 template: (dz, dx) = (-1,-1), (-1,0), (-1,1), (0,-1)
    assert list(offs) == [nx + 1, nx, nx - 1, 1]
    mask5 = psvi_mask_2D(np.ones((5, 5)), 400 * 200, nx=400, nz=200)
    assert mask5.shape == (12, 80000)
    assert sorted((mask5 == 0).sum(axis=1)) == sorted([2 * 400 + d for d in (-2, -1, 0, 1, 2)]
                                                       + [400 + d for d in (-2, -1, 0, 1, 2)] + [1, 2])


def test_structured_oracle_is_banded_lower_triangular_matvec():
    dim, offs = 9, [1, 3, 4]
    u, dl = RNG.normal(size=dim), RNG.normal(size=dim) * 0.3
    nd = RNG.normal(size=(3, dim))
    x = RNG.normal(size=(4, dim))
    L = O.structured_dense(dl, nd, offs)
    assert np.allclose(np.triu(L, 1), 0)
    z, ld = O.linear_forward(u, dl, x.copy(), "structured", nd, offs)
    assert np.allclose(z, u + x @ L.T, atol=1e-13)
    assert np.allclose(ld, np.linalg.slogdet(L)[1], atol=1e-12)


def test_logistic_jacobian_matches_finite_differences():
    lo, up = np.array([1400.0, 1.0]), np.array([4800.0, 2.0])
    x = RNG.normal(size=(3, 2))
    _, ld = O.real2constr(lo, up, x)
    h = 1e-6
    jac = np.zeros_like(x)
    for j in range(2):
        e = np.zeros(2)
        e[j] = h
        jac[:, j] = (O.real2constr(lo, up, x + e)[0][:, j] - O.real2constr(lo, up, x - e)[0][:, j]) / (2 * h)
    assert np.allclose(ld, np.log(jac).sum(axis=1), rtol=1e-7)
    z = O.real2constr(lo, up, x)[0]
    xb, ld_inv = O.constr2real(lo, up, z)
    assert np.allclose(xb, x, atol=1e-9)
    assert np.allclose(ld_inv, -ld, atol=1e-9)


def test_smoothing_matrix_annihilates_constant_fields():
    nx, nz = 6, 4
    S = smooth_matrix_2D(nx, nz, 3.0, 2.0)
    assert S.shape == (2 * nx * nz, nx * nz)
    assert np.allclose(S @ np.full(nx * nz, 7.0), 0)
    # a field linear in x has zero second differences away from the two ends
    f = np.tile(np.arange(nx, dtype=float), nz)
    r = (S @ f)[:nx * nz].reshape(nz, nx)
    assert np.allclose(r[:, 1:-1], 0) and np.allclose(r[:, 0] * 3.0, 1) and np.allclose(r[:, -1] * 3.0, -1)


# ---- torch modules (CPU) vs oracle ------------------------------------------

@pytest.mark.parametrize("kernel", ["diagonal", "structured", "fullrank"])
def test_linear_matches_oracle_and_inverts(kernel):
    dim = 12
    mask = psvi_mask_2D(np.ones((3, 3)), dim, nx=4, nz=3)
    f = Linear(dim, kernel=kernel, mask=mask)
    with torch.no_grad():
        f.u.copy_(torch.from_numpy(RNG.normal(size=dim)))
        f.diag.copy_(torch.from_numpy(RNG.normal(size=dim) * 0.2))
        if f.non_diag is not None:
            f.non_diag.copy_(torch.from_numpy(RNG.normal(size=tuple(f.non_diag.shape)) * 0.1))
    x = RNG.normal(size=(5, dim))
    z, ld = f(torch.from_numpy(x))
    nd = f.non_diag.detach().numpy() if f.non_diag is not None else None
    zo, ldo = O.linear_forward(f.u.detach().numpy(), f.diag.detach().numpy(), x.copy(), kernel, nd,
                               list((mask == 0).sum(axis=1)))
    assert np.allclose(z.detach().numpy(), zo, atol=1e-12)
    assert np.allclose(ld.detach().numpy(), ldo, atol=1e-12)
    xb, ldi = f.inverse(z)
    assert np.allclose(xb.detach().numpy(), x, atol=1e-10)
    assert np.allclose(ldi.detach().numpy(), -ldo, atol=1e-12)


def test_param_initialisation_layout():
    dim = 6
    mask = psvi_mask_2D(np.ones((3, 3)), dim, nx=3, nz=2)
    p = np.concatenate([np.arange(dim) + 1.0, np.full(dim, 2.0), np.full(mask.size, 0.5)])
    f = Linear(dim, kernel="structured", mask=mask, param=p)
    assert np.allclose(f.u.detach().numpy(), np.arange(dim) + 1.0)
    assert np.allclose(np.exp(f.diag.detach().numpy()), 2.0)
    assert np.allclose(f.non_diag.detach().numpy(), 0.5)


def test_constraint_flows_and_uniform_prior_match_oracle():
    dim = 8
    lo, up = np.full(dim, 1400.0), np.full(dim, 4800.0)
    x = RNG.normal(size=(3, dim))
    z, ld = Real2Constr(dim, lower=lo, upper=up)(torch.from_numpy(x))
    zo, ldo = O.real2constr(lo, up, x)
    assert np.allclose(z.numpy(), zo) and np.allclose(ld.numpy(), ldo)
    xb, ldb = Constr2Real(dim, lower=lo, upper=up)(z)
    assert np.allclose(xb.numpy(), x, atol=1e-9) and np.allclose(ldb.numpy(), -ldo, atol=1e-9)
    S = smooth_matrix_2D(4, 2, 5.0, 5.0)
    pr = Uniform(lower=lo, upper=up, smooth_matrix=S)
    zt = torch.from_numpy(zo).requires_grad_()
    lp = pr.log_prob(zt)
    assert np.allclose(lp.detach().numpy(), O.uniform_log_prob(lo, up, zo, S))
    lp.sum().backward()
    # gradient of -0.5||Sx||^2 is -S^T S x
    assert np.allclose(zt.grad.numpy(), -(S.T @ (S @ zo.T)).T)
    assert np.allclose(Uniform(lower=lo, upper=up).log_prob(zt).item(), O.uniform_log_prob(lo, up, zo)[0])


def test_flows_based_distribution_density_and_seeded_base():
    dim = 10
    lo, up = np.full(dim, 1.0), np.full(dim, 3.0)
    flows = [Linear(dim, "structured", mask=psvi_mask_2D(np.ones((3, 3)), dim, nx=5, nz=2)),
             Real2Constr(dim, lower=lo, upper=up)]
    with torch.no_grad():
        flows[0].non_diag.normal_(0, 0.1)
        flows[0].diag.normal_(0, 0.1)
    q = FlowsBasedDistribution(flows, base="Normal", generator=torch.Generator().manual_seed(3))
    q.double()
    x = q.sample_from_base(6)
    x2 = FlowsBasedDistribution(flows, generator=torch.Generator().manual_seed(3)).sample_from_base(6)
    assert torch.equal(x, x2)
    z, logq = q(x)
    xb, logq_inv = q.inverse(z)
    assert np.allclose(xb.detach().numpy(), x.numpy(), atol=1e-9)
    assert np.allclose(logq.detach().numpy(), logq_inv.detach().numpy(), atol=1e-8)
    # logq = log q0(x) - sum of log|det|
    ld = flows[0](x)[1] + flows[1](flows[0](x)[0])[1]
    assert np.allclose(logq.detach().numpy(), (O.log_prob_base(x.numpy()) - ld.detach().numpy()), atol=1e-10)
