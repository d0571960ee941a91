"""Pins of the oracle's adjoint and injection terms by identities the
mathematics fixes (no retyping of the oracle's formulas).

* Source-receiver reciprocity of the records.  The scheme is
  M (u^{n+1} - 2u^n + u^{n-1}) = dt^2 (L u^n + f^n delta_s) with M = diag(1/v^2)
  and L the symmetric finite-difference Laplacian, so the record at r of a
  source at s equals the record at s of a source at r, exactly, until the
  absorbing boundary is reached -- if and only if the source enters as
  rik*dt^2*v_s^2 (fdtd2D_modelling.c:910).
* Adjoint = time-reversed forward.  With the residual at a single receiver R
  set to the time-reversed wavelet, the adjoint field of backward_aco_2D
  (residual injected as res*dt^2*v_R^2, fdtd2D_modelling.c:955) is the
  forward field of a source at R, reversed in time: lambda(it) = u_R(nt-1-it).
  The gradient is then a closed form in two forward fields, built here from
  the oracle's single-step propagator (pinned in test_oracle_pins):
    run_fwi=1: sum_{it%k==0} -2/v^3 * u_tt,S(it) * u_R(nt-1-it) * k * dt
    run_fwi=0: sum_{it%k==0} u_S(it) * u_R(nt-1-it) * k * dt
  (fdtd2D_modelling.c:913-924, 959-975, with u_tt the stored
  (p2 - 2 p1 + p0)/dt/dt and residual = syn - obs or obs, :272-277).
* The same identity swaps source and receiver: the gradient of (S, R) equals
  the gradient of (R, S), for both run_fwi modes (summation by parts in time
  for u_tt).
* The imaging interval: sum_{it%k==0} (...) * k * dt is a Riemann sum of a
  band-limited integrand, so the gradient at k = 2 and 3 equals the k = 1
  gradient to ~1e-6; a dropped or doubled factor k, or a shifted store slot
  (it/k, fdtd2D_modelling.c:917, 969-974), moves it by 50 % or more.

The check_* functions raise AssertionError; tests/test_oracle_mutants.py runs
them against deliberately mutated oracle builds and requires each mutation
to fail at least one of them.
"""
import numpy as np

from geopvi_amd import synth
from geopvi_amd.forward.fwi2d.params import FWIParams
from helpers import rel_l2

NZ, NX = 60, 80
A = (22, 18)   # (depth, column) in inner-grid cells; v ~ 1840 m/s
B = (38, 60)   # v ~ 2220 m/s


def hetero_model():
    """Laterally and vertically varying model: v(A) != v(B)."""
    zz, xx = np.meshgrid(np.arange(NZ), np.arange(NX), indexing="ij")
    return 1500.0 + 1500.0 * xx / (NX - 1) + 600.0 * np.sin(zz / 7.0) * np.cos(xx / 9.0)


def point_params(src, rec, nt, k=1, Lc=3):
    """One shot at src, one receiver at rec (inner-grid (iz, ix))."""
    return FWIParams(nx=NX, nz=NZ, pml0=10, Lc=Lc, laplace_slover=0, ns=1, nt=nt, ds=1, ns0=src[1],
                     depths=src[0], depthr=rec[0], nr=1, dr=1, nr0=rec[1], nt_interval=k, dx=10.0, dz=10.0,
                     dt=0.001, f0=30.0)


def forward_fields(oracle_mod, p, v, src, dtype=np.float64):
    """Every forward field u(it) (padded, (nt, ntz, ntx)) of a source at src,
    from the oracle's single-step propagator plus the source term."""
    velp = oracle_mod.pad_velocity(p, v, dtype)
    rik = oracle_mod.ricker(p.nt, p.dt, p.f0, dtype)
    dt = dtype(np.float32(p.dt))
    sz, sx = p.pml + src[0], p.pml + src[1]
    p0 = np.zeros((p.ntz, p.ntx), dtype)
    p1 = np.zeros_like(p0)
    out = np.zeros((p.nt, p.ntz, p.ntx), dtype)
    for it in range(p.nt):
        p2 = oracle_mod.step(p, velp, p0, p1, dtype)
        p2[sz, sx] += rik[it] * dt * dt * velp[sz, sx] * velp[sz, sx]
        out[it] = p2
        p0, p1 = p1, p2
    return out


def reversed_residual_obs(oracle_mod, p, v, run_fwi, dtype):
    """Observed data whose residual is the time-reversed wavelet."""
    rev = oracle_mod.ricker(p.nt, p.dt, p.f0, dtype)[::-1].copy()
    if not run_fwi:
        return rev
    syn = oracle_mod.forward(p, v, dtype=dtype)
    return (syn - rev).astype(dtype)


def check_reciprocity_records(oracle_mod, dtype=np.float64, tol=1e-10, nt=250):
    v = hetero_model()
    r_ab = oracle_mod.forward(point_params(A, B, nt), v, dtype=dtype)
    r_ba = oracle_mod.forward(point_params(B, A, nt), v, dtype=dtype)
    assert np.abs(r_ab).max() > 1e-3, "the wave must reach the receiver"
    err = rel_l2(r_ab, r_ba)
    assert err < tol, f"records A->B vs B->A: rel {err:.3e}"
    return err


def check_adjoint_closed_form(oracle_mod, run_fwi, k=1, nt=251, tol=1e-10):
    v = hetero_model()
    p = point_params(A, B, nt, k)
    obs = reversed_residual_obs(oracle_mod, p, v, run_fwi, np.float64)
    _, g = oracle_mod.fwi(p, v, obs, run_fwi=run_fwi, dtype=np.float64)
    uS = forward_fields(oracle_mod, p, v, A)
    uR = forward_fields(oracle_mod, p, v, B)
    dt = float(np.float32(p.dt))
    inner = (slice(None), slice(p.pml, p.ntz - p.pml), slice(p.pml, p.ntx - p.pml))
    uS, uR = uS[inner], uR[inner]
    vin = oracle_mod.pad_velocity(p, v, np.float64)[inner[1:]]
    expect = np.zeros((NZ, NX))
    for it in range(0, nt, k):
        lam = uR[nt - 1 - it]
        if run_fwi:
            prev1 = uS[it - 1] if it >= 1 else 0.0
            prev2 = uS[it - 2] if it >= 2 else 0.0
            utt = (uS[it] - 2 * prev1 + prev2) / dt / dt
            expect += -2.0 / vin ** 3 * utt * lam * k * dt
        else:
            expect += uS[it] * lam * k * dt
    assert np.linalg.norm(expect) > 0
    err = rel_l2(g, expect)
    assert err < tol, f"run_fwi={run_fwi} k={k}: gradient vs closed form rel {err:.3e}"
    return err


def check_gradient_swap(oracle_mod, run_fwi, dtype=np.float64, tol=1e-10, nt=251):
    v = hetero_model()
    gs = []
    for s, r in ((A, B), (B, A)):
        p = point_params(s, r, nt)
        obs = reversed_residual_obs(oracle_mod, p, v, run_fwi, dtype)
        gs.append(oracle_mod.fwi(p, v, obs, run_fwi=run_fwi, dtype=dtype)[1])
    assert np.linalg.norm(gs[0]) > 0
    err = rel_l2(gs[0], gs[1])
    assert err < tol, f"run_fwi={run_fwi}: gradient (S, R) vs (R, S) rel {err:.3e}"
    return err


def interval_case():
    p = synth.workload_params("tiny")
    vt = synth.gradient_anomaly_model(p.nz, p.nx, seed=2)
    return p, vt, synth.smooth(vt, 3.0)


def check_interval_invariance(oracle_mod, run_fwi, tol=1e-3):
    p, vt, vs = interval_case()
    obs = oracle_mod.forward(p, vt, dtype=np.float64)
    g = {k: oracle_mod.fwi(p.replace(nt_interval=k), vs, obs, run_fwi=run_fwi, dtype=np.float64)[1]
         for k in (1, 2, 3)}
    errs = {k: rel_l2(g[k], g[1]) for k in (2, 3)}
    for k, e in errs.items():
        assert e < tol, f"run_fwi={run_fwi}: gradient at k={k} vs k=1 rel {e:.3e}"
    return errs


# ---------------------------------------------------------------------------

def test_reciprocity_records_f64(oracle_mod):
    print("rel", check_reciprocity_records(oracle_mod))


def test_reciprocity_records_f32(oracle_mod):
    print("rel", check_reciprocity_records(oracle_mod, np.float32, tol=1e-5))


def test_reciprocity_breaks_once_the_boundary_is_reached(oracle_mod):
    # the absorbing boundary is not symmetric: the identity is about the
    # interior scheme, and the test above stops before the boundary matters
    v = hetero_model()
    r_ab = oracle_mod.forward(point_params(A, B, 420), v)
    r_ba = oracle_mod.forward(point_params(B, A, 420), v)
    assert rel_l2(r_ab, r_ba) > 1e-4


def test_adjoint_closed_form_run_fwi_1(oracle_mod):
    print("rel", check_adjoint_closed_form(oracle_mod, 1))


def test_adjoint_closed_form_run_fwi_0(oracle_mod):
    print("rel", check_adjoint_closed_form(oracle_mod, 0))


def test_adjoint_closed_form_k2(oracle_mod):
    # nt - 1 even, so every imaging step it has a partner nt - 1 - it
    for run_fwi in (0, 1):
        print("rel", check_adjoint_closed_form(oracle_mod, run_fwi, k=2))


def test_gradient_swap_symmetry(oracle_mod):
    for run_fwi in (0, 1):
        print("f64 rel", check_gradient_swap(oracle_mod, run_fwi))
        print("f32 rel", check_gradient_swap(oracle_mod, run_fwi, np.float32, tol=5e-6))


def test_imaging_interval_invariance(oracle_mod):
    for run_fwi in (0, 1):
        print(check_interval_invariance(oracle_mod, run_fwi))


def test_run_fwi_0_gradient_bilinear_and_zero(oracle_mod):
    p, vt, vs = interval_case()
    obs = oracle_mod.forward(p, vt, dtype=np.float64)
    _, g1 = oracle_mod.fwi(p, vs, obs, run_fwi=0, dtype=np.float64)
    _, g2 = oracle_mod.fwi(p, vs, 3.0 * obs, run_fwi=0, dtype=np.float64)
    syn0, g0 = oracle_mod.fwi(p, vs, np.zeros_like(obs), run_fwi=0, dtype=np.float64)
    assert np.linalg.norm(g1) > 0
    assert rel_l2(g2, 3.0 * g1) < 1e-12
    assert not g0.any()
    # the records of run_fwi=0 are the forward records
    assert np.array_equal(syn0, oracle_mod.forward(p, vs, dtype=np.float64))


def test_row_major_build_gives_the_same_bits(oracle_mod):
    """The oracle's row-major build (used for long fixture runs) computes the
    same arithmetic in another loop order: identical records and gradients."""
    import oracle
    p, vt, vs = interval_case()
    obs = oracle.forward(p, vt)
    prev = oracle.select(row_major=True)
    try:
        rm = oracle.fwi(p, vs, obs), oracle.fwi(p, vs, obs, run_fwi=0)
    finally:
        oracle.use_library(prev)
    ref = oracle.fwi(p, vs, obs), oracle.fwi(p, vs, obs, run_fwi=0)
    for (s1, g1), (s2, g2) in zip(rm, ref):
        assert np.array_equal(s1, s2) and np.array_equal(g1, g2)
