"""Shared helpers for the parity tests: seeded cases and error measures."""
from __future__ import annotations

import numpy as np

from geopvi_amd import synth


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def rel_err(a: float, b: float) -> float:
    return abs(float(a) - float(b)) / max(abs(float(b)), 1e-300)


def case(name: str, seed: int = 0, smooth_sigma: float = None, **overrides):
    """(params, true model, starting model) for a named workload."""
    p = synth.workload_params(name, **overrides)
    vt = synth.gradient_anomaly_model(p.nz, p.nx, p.dx, p.dz, seed=seed)
    sig = smooth_sigma if smooth_sigma is not None else max(1.0, min(p.nz, p.nx) / 15.0)
    vs = synth.smooth(vt, sig)
    return p, vt, vs


# Tolerances stated by the parity contract (FP32): misfit relative error
# <= 1e-5, gradient relative L2 <= 1e-4.
MISFIT_RTOL = 1e-5
GRAD_RTOL = 1e-4


def fwi_config(p):
    """configparser object with the [FWI] section the reference's Posterior
    reads (examples/fwi2d/input/config.ini keys) for run parameters p."""
    import configparser
    cfg = configparser.ConfigParser()
    keys = [("nx", "nx"), ("nz", "nz"), ("pml0", "pml"), ("Lc", "Lc"), ("laplace_slover", "laplace_slover"),
            ("ns", "ns"), ("nt", "nt"), ("ds", "ds"), ("ns0", "ns0"), ("depths", "depths"), ("depthr", "depthr"),
            ("nr", "nr"), ("dr", "dr"), ("nr0", "nr0"), ("nt_interval", "nt_interval"), ("dx", "dx"),
            ("dz", "dz"), ("dt", "dt"), ("f0", "f0")]
    cfg["FWI"] = {key: (repr(getattr(p, name)) if name in ("dx", "dz", "dt", "f0") else str(getattr(p, name)))
                  for name, key in keys}
    return cfg
