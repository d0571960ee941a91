"""Invariants of the full oracle solve (forward records and gradient).

* left-right mirror symmetry of a laterally uniform model (bit-exact),
* moveout of the direct wave = offset / v in a homogeneous medium,
* zero residual -> zero gradient; gradient linear in the residual,
* the multi-shot gradient is the sum of single-shot gradients,
* the float build agrees with the double twin,
* the reference gradient against central finite differences of the misfit
  (the reference uses a continuous adjoint, so agreement is approximate).
"""
import numpy as np
import pytest

from geopvi_amd import synth
from geopvi_amd.forward.fwi2d.params import FWIParams
from helpers import rel_l2


def tiny_params(**kw):
    base = dict(nx=40, nz=24, pml0=10, Lc=3, ns=1, ds=1, ns0=20, depths=0, depthr=0, nr=41, dr=1,
                nr0=-0, nt=260, nt_interval=2, dx=10.0, dz=10.0, dt=0.001, f0=25.0)
    base.update(kw)
    return FWIParams(**base)


def test_mirror_symmetry_bit_exact(oracle_mod):
    # source at the centre column of a laterally uniform model, receivers at
    # every column: record(ir) == record(nr-1-ir) exactly.
    p = tiny_params(nx=41, ns0=20, nr=41)
    v = np.broadcast_to(np.linspace(1500, 2500, p.nz)[:, None], (p.nz, p.nx)).astype(np.float32)
    rec = oracle_mod.forward(p, v).reshape(p.nt, p.nr)
    assert np.abs(rec).max() > 0
    assert np.array_equal(rec, rec[:, ::-1])


def test_direct_wave_moveout(oracle_mod):
    # homogeneous 2000 m/s; lag between traces at offsets 100 m and 300 m is
    # 200 m / 2000 m/s = 0.1 s = 100 steps (within 2 steps of dispersion)
    p = tiny_params(nx=120, nz=40, ns0=10, nr=120, nt=600, f0=20.0, depths=5, depthr=5)
    v = np.full((p.nz, p.nx), 2000.0, np.float32)
    rec = oracle_mod.forward(p, v, dtype=np.float64).reshape(p.nt, p.nr)
    a, b = rec[:, 20], rec[:, 40]   # offsets 10 and 30 cells
    lags = np.arange(-300, 301)
    cc = [np.dot(a[max(0, -l):p.nt - max(0, l)], b[max(0, l):p.nt - max(0, -l)]) for l in lags]
    lag = lags[int(np.argmax(cc))]
    assert abs(lag - 100) <= 2, lag


def test_zero_residual_gives_zero_gradient(oracle_mod):
    p = synth.workload_params("tiny")
    v = synth.gradient_anomaly_model(p.nz, p.nx, seed=2).astype(np.float32)
    obs = oracle_mod.forward(p, v)
    syn, grad = oracle_mod.fwi(p, v, obs)
    assert np.array_equal(syn, obs)
    assert not grad.any()


def test_gradient_linear_in_residual(oracle_mod):
    p = synth.workload_params("tiny")
    vt = synth.gradient_anomaly_model(p.nz, p.nx, seed=2)
    vs = synth.smooth(vt, 3.0)
    obs = oracle_mod.forward(p, vt, dtype=np.float64)
    syn, g1 = oracle_mod.fwi(p, vs, obs, dtype=np.float64)
    obs2 = syn - 2.0 * (syn - obs)          # residual doubled
    _, g2 = oracle_mod.fwi(p, vs, obs2, dtype=np.float64)
    assert rel_l2(g2, 2.0 * g1) < 1e-12


def test_shot_additivity(oracle_mod):
    p = synth.workload_params("tiny")
    vt = synth.gradient_anomaly_model(p.nz, p.nx, seed=4)
    vs = synth.smooth(vt, 3.0)
    obs = oracle_mod.forward(p, vt, dtype=np.float64)
    _, g_all = oracle_mod.fwi(p, vs, obs, dtype=np.float64)
    parts = [oracle_mod.fwi(p, vs, obs, dtype=np.float64, shots=(s, s + 1))[1] for s in range(p.ns)]
    assert rel_l2(sum(parts), g_all) < 1e-12


def test_float_build_matches_double_twin(oracle_mod):
    p = synth.workload_params("small")
    vt = synth.gradient_anomaly_model(p.nz, p.nx, seed=5)
    vs = synth.smooth(vt, 4.0)
    obs64 = oracle_mod.forward(p, vt, dtype=np.float64)
    syn64, g64 = oracle_mod.fwi(p, vs, obs64, dtype=np.float64)
    syn32, g32 = oracle_mod.fwi(p, vs, obs64.astype(np.float32))
    assert rel_l2(syn32, syn64) < 1e-5
    m32 = 0.5 * np.sum((syn32.astype(np.float64) - obs64.astype(np.float32)) ** 2)
    m64 = 0.5 * np.sum((syn64 - obs64.astype(np.float32).astype(np.float64)) ** 2)
    assert abs(m32 - m64) / m64 < 1e-5
    assert rel_l2(g32, g64) < 1e-4


def test_gradient_against_finite_differences(oracle_mod):
    """Directional derivatives of the misfit vs <-grad, d> (posterior.py:110
    returns -grad as the misfit gradient).

    The reference runs the forward propagator backwards with the residual
    injected as res*dt^2*v_r^2 (fdtd.c:955) and images -2/v^3 * u_tt * q *
    k * dt (fdtd.c:965).  With u^{n+1} = 2u^n - u^{n-1} + dt^2 V L u^n the
    exact adjoint lambda satisfies the same recursion for mu = V lambda, so
    q = dt^2 v^2 lambda and the exact derivative sum_n 2 dt^2/v u_tt lambda
    equals -grad / dt: one scale, 1/dt, in the interior.  Two parts of the
    discrete misfit are outside the image: the ABC, which is not adjointed,
    and the absorbing bands' velocities, which are edge-replicated copies of
    the model's outermost rows and columns (fdtd.c:1197-1238) -- a
    perturbation of an edge cell moves a whole line of band cells next to the
    surface sources and receivers.  So the directions vanish on the model's
    edge, and after the best positive scale the directional derivatives
    agree within 5 %, their cosine exceeds 0.99, and the scale lies in
    [0.7/dt, 1.4/dt] (0.77/dt measured), a window that a factor 2 either way
    leaves.  The imaging interval does not change the scale: at
    nt_interval = 2 it is the k = 1 scale within 1 % (a dropped or doubled
    factor k, fdtd2D_modelling.c:965, would halve or double it)."""
    p = synth.workload_params("tiny", nt_interval=1)
    vt = synth.gradient_anomaly_model(p.nz, p.nx, seed=6)
    vs = synth.smooth(vt, 3.0)
    obs = oracle_mod.forward(p, vt, dtype=np.float64)
    _, g = oracle_mod.fwi(p, vs, obs, dtype=np.float64)
    _, g2 = oracle_mod.fwi(p.replace(nt_interval=2), vs, obs, dtype=np.float64)
    grad_misfit = -g

    def misfit(v):
        s = oracle_mod.forward(p, v, dtype=np.float64)
        return 0.5 * np.sum((s - obs) ** 2)

    rng = np.random.default_rng(11)
    fd, an, an2 = [], [], []
    for _ in range(4):
        d = synth.smooth(rng.standard_normal((p.nz, p.nx)), 2.0) * synth.edge_taper(p.nz, p.nx, 3)
        d /= np.abs(d).max()
        eps = 1.0  # m/s
        fd.append((misfit(vs + eps * d) - misfit(vs - eps * d)) / (2 * eps))
        an.append(float(np.dot(grad_misfit, d.reshape(-1))))
        an2.append(float(np.dot(-g2, d.reshape(-1))))
    fd, an, an2 = np.array(fd), np.array(an), np.array(an2)
    cos = float(np.dot(fd, an) / (np.linalg.norm(fd) * np.linalg.norm(an)))
    alpha = float(np.dot(fd, an) / np.dot(an, an))
    alpha2 = float(np.dot(fd, an2) / np.dot(an2, an2))
    rel = float(np.linalg.norm(fd - alpha * an) / np.linalg.norm(fd))
    print(f"FD check: cos={cos:.5f} scale={alpha:.1f} (k=2: {alpha2:.1f}) rel-after-scale={rel:.3e}")
    dt = float(np.float32(p.dt))
    assert cos > 0.99 and rel < 0.05
    assert 0.7 / dt < alpha < 1.4 / dt
    assert abs(alpha2 / alpha - 1.0) < 0.01
