"""Seeded synthetic velocity models and acquisition set-ups.

All arrays come from ``numpy.random.default_rng(seed)`` (PCG64), so every
machine generates identical inputs.  Models are ``(nz, nx)`` float64 in m/s,
row-major with x fastest, the layout the solver's ``vel_inner`` expects
(geopvi/forward/fwi2d/fdtd2D_modelling.c:1199-1203).

Workloads (grid nz x nx at dx = dz = 10 m, dt = 1 ms, pml0 = 10, Lc = 3,
nt_interval = 2, surface sources and receivers at grid row 0):

* ``c0``: 120 x 200, 4 shots, 100 receivers, 10 Hz Ricker, 1000 steps;
  true model = linear 1500 -> 3500 m/s gradient + 3 Gaussian anomalies of
  +-300 m/s (radius 150 m); the sample is the truth smoothed (sigma 8 cells).
* ``c1``: 200 x 400, 20 shots, 200 receivers, 8 Hz, 2000 steps; 16 samples
  = smoothed truth + N(0, 50 m/s) per cell; observed data + 1 % RMS noise.
* ``c2``: 300 x 900, 30 shots, 300 receivers, 10 Hz, 4000 steps; 8 layers
  1500 -> 4500 m/s cut by 2 dipping faults (200 m throw) and a 400 m wide
  low-velocity lens (-400 m/s).
"""
from __future__ import annotations

import math

import numpy as np

from .forward.fwi2d.params import FWIParams

_COMMON = dict(pml0=10, Lc=3, laplace_slover=0, depths=0, depthr=0, nt_interval=2,
               dx=10.0, dz=10.0, dt=0.001)

WORKLOADS = {
    "c0": dict(nz=120, nx=200, ns=4, ds=50, ns0=25, nr=100, dr=2, nr0=0, f0=10.0, nt=1000),
    "c1": dict(nz=200, nx=400, ns=20, ds=20, ns0=10, nr=200, dr=2, nr0=0, f0=8.0, nt=2000),
    "c2": dict(nz=300, nx=900, ns=30, ds=30, ns0=15, nr=300, dr=3, nr0=0, f0=10.0, nt=4000),
    # small cases for parity tests
    "tiny": dict(nz=24, nx=40, ns=2, ds=18, ns0=3, nr=20, dr=2, nr0=0, f0=25.0, nt=300),
    "small": dict(nz=40, nx=64, ns=3, ds=20, ns0=10, nr=32, dr=2, nr0=0, f0=20.0, nt=400),
}


def workload_params(name: str, **overrides) -> FWIParams:
    kw = dict(_COMMON)
    kw.update(WORKLOADS[name])
    kw.update(overrides)
    return FWIParams(**kw)


def gradient_anomaly_model(nz: int, nx: int, dx: float = 10.0, dz: float = 10.0, seed: int = 0,
                           v_top: float = 1500.0, v_bottom: float = 3500.0, n_anom: int = 3,
                           amp: float = 300.0, radius: float = 150.0,
                           margin: float = 150.0) -> np.ndarray:
    """Linear depth gradient plus ``n_anom`` Gaussian anomalies of +-amp."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz, dtype=np.float64)[:, None] * dz
    x = np.arange(nx, dtype=np.float64)[None, :] * dx
    v = v_top + (v_bottom - v_top) * (np.arange(nz, dtype=np.float64)[:, None] / max(nz - 1, 1))
    v = np.broadcast_to(v, (nz, nx)).copy()
    zmax, xmax = (nz - 1) * dz, (nx - 1) * dx
    for _ in range(n_anom):
        sign = 1.0 if rng.random() < 0.5 else -1.0
        cz = rng.uniform(min(margin, zmax / 2), max(zmax - margin, zmax / 2))
        cx = rng.uniform(min(margin, xmax / 2), max(xmax - margin, xmax / 2))
        v += sign * amp * np.exp(-((z - cz) ** 2 + (x - cx) ** 2) / radius ** 2)
    return v


def smooth(v: np.ndarray, sigma: float = 8.0) -> np.ndarray:
    """Separable Gaussian smoothing with edge replication (truncated at 4 sigma)."""
    r = int(4.0 * sigma + 0.5)
    t = np.arange(-r, r + 1, dtype=np.float64)
    k = np.exp(-0.5 * (t / sigma) ** 2)
    k /= k.sum()
    out = v.astype(np.float64)
    for axis in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[axis] = (r, r)
        p = np.pad(out, pad, mode="edge")
        acc = np.zeros_like(out)
        n = out.shape[axis]
        for j, kj in enumerate(k):
            sl = [slice(None), slice(None)]
            sl[axis] = slice(j, j + n)
            acc += kj * p[tuple(sl)]
        out = acc
    return out


def layered_fault_lens_model(nz: int = 300, nx: int = 900, dx: float = 10.0, dz: float = 10.0,
                             seed: int = 2, n_layers: int = 8, v_min: float = 1500.0,
                             v_max: float = 4500.0, throw: float = 200.0,
                             lens_width: float = 400.0, lens_dv: float = -400.0) -> np.ndarray:
    """8 flat layers offset by two dipping normal faults, plus a lens.

    Fault k crosses the surface at x_k and dips at angle a_k; every interface
    to the right of fault k is shifted down by ``throw``.  The lens is an
    ellipse ``lens_width`` wide and a quarter of that tall.
    """
    rng = np.random.default_rng(seed)
    z = np.arange(nz, dtype=np.float64)[:, None] * dz
    x = np.arange(nx, dtype=np.float64)[None, :] * dx
    depth = (nz - 1) * dz
    width = (nx - 1) * dx
    vels = np.linspace(v_min, v_max, n_layers)
    tops = np.linspace(0.0, depth, n_layers + 1)[1:-1]           # 7 interfaces
    shift = np.zeros((nz, nx))
    for k in range(2):
        x0 = width * (0.3 + 0.4 * k) + rng.uniform(-0.05, 0.05) * width
        dip = math.radians(rng.uniform(55.0, 70.0))
        xf = x0 + z / math.tan(dip)
        shift += np.where(x > xf, throw, 0.0)
    zz = z - shift                                                # depth in the unfaulted frame
    layer = np.zeros((nz, nx), dtype=np.int64)
    for t in tops:
        layer += (zz >= t).astype(np.int64)
    v = vels[np.clip(layer, 0, n_layers - 1)]
    cz = rng.uniform(0.35, 0.6) * depth
    cx = rng.uniform(0.3, 0.7) * width
    a, b = lens_width / 2.0, lens_width / 8.0
    inside = ((x - cx) / a) ** 2 + ((z - cz) / b) ** 2 <= 1.0
    v = np.where(inside, v + lens_dv, v)
    return v


def gaussian_samples(mean: np.ndarray, std: float, n: int, seed: int) -> np.ndarray:
    """``n`` draws of a mean-field Gaussian, shape (n, nz*nx)."""
    rng = np.random.default_rng(seed)
    flat = mean.reshape(-1)
    return flat[None, :] + std * rng.standard_normal((n, flat.size))


def edge_taper(nz: int, nx: int, width: int) -> np.ndarray:
    """(nz, nx) weights: 0 on the outermost row and column, rising as
    sin^2 to 1 over ``width`` cells, 1 inside.  Multiplying a model
    perturbation by it leaves the edge-replicated absorbing bands unchanged."""
    def t(n):
        w = np.ones(n)
        r = np.sin(0.5 * np.pi * np.arange(width) / width) ** 2
        w[:width] = r
        w[n - width:] = np.minimum(w[n - width:], r[::-1])
        return w
    return t(nz)[:, None] * t(nx)[None, :]


def add_noise(data: np.ndarray, frac: float, seed: int) -> np.ndarray:
    """Additive Gaussian noise with std = frac * RMS(data)."""
    rng = np.random.default_rng(seed)
    rms = float(np.sqrt(np.mean(np.asarray(data, np.float64) ** 2)))
    return (data + frac * rms * rng.standard_normal(data.shape)).astype(data.dtype)


def true_model(name: str, seed: int = 0) -> np.ndarray:
    p = workload_params(name)
    if name == "c2":
        return layered_fault_lens_model(p.nz, p.nx, p.dx, p.dz, seed=seed)
    return gradient_anomaly_model(p.nz, p.nx, p.dx, p.dz, seed=seed)


SERVING_MIXES = {
    # grids, nsample range, shot range, step range (inclusive)
    "c3": ([(100, 200), (200, 400), (300, 600), (400, 800)], (4, 64), (5, 40), (1000, 3000)),
    # the same generator at a scale whose 200-request stream takes seconds
    "bench": ([(100, 200), (200, 400)], (4, 16), (5, 10), (500, 1000)),
}


def serving_request(rng: np.random.Generator, mix: str = "c3"):
    """One request of a serving mix: grid, samples, shots and steps drawn
    uniformly from the mix's sets (c3: {100x200, 200x400, 300x600,
    400x800}, 4..64, 5..40 and 1000..3000); receivers every 2 cells across
    half the width, shots spread evenly."""
    grids, (m0, m1), (s0, s1), (t0, t1) = SERVING_MIXES[mix]
    nz, nx = grids[int(rng.integers(0, len(grids)))]
    nsample = int(rng.integers(m0, m1 + 1))
    ns = int(rng.integers(s0, s1 + 1))
    nt = int(rng.integers(t0, t1 + 1))
    ds = nx // ns
    kw = dict(_COMMON)
    kw.update(nz=nz, nx=nx, ns=ns, ds=ds, ns0=ds // 2, nr=nx // 2, dr=2, nr0=0, f0=10.0, nt=nt)
    return FWIParams(**kw), nsample
