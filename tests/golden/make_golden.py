#!/usr/bin/env python3
"""Write the golden fixtures of the FWI-2D path.

The reference ships no test vectors for this path (no tests, and the
example's waveform.npy is not in the repository), so the fixtures are the
CPU oracle's outputs on small seeded cases; the oracle itself is pinned by
tests/test_oracle_pins.py and tests/test_oracle_physics.py.  Each fixture
holds the 19 run parameters, the true and starting models, the observed
records (forward of the true model) and the float-build records, misfit and
gradient of the starting model, plus the double-twin gradient.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)

import oracle  # noqa: E402
from geopvi_amd import synth  # noqa: E402
from geopvi_amd.forward.fwi2d.params import FWIParams  # noqa: E402

CASES = {
    "tiny_k2": dict(workload="tiny", seed=3, smooth=2.0, over={}),
    "tiny_k1_lc6": dict(workload="tiny", seed=5, smooth=2.0, over=dict(Lc=6, nt=200, nt_interval=1)),
    "small_k3": dict(workload="small", seed=7, smooth=3.0, over=dict(nt_interval=3, depths=2, depthr=1)),
}


def param_vector(p: FWIParams) -> np.ndarray:
    return np.array([getattr(p, n) for n, _ in oracle.OParams._fields_], dtype=np.float64)


def params_from_vector(v) -> FWIParams:
    names = [n for n, _ in oracle.OParams._fields_]
    kw = {n: (int(x) if i < 15 else float(x)) for i, (n, x) in enumerate(zip(names, v))}
    return FWIParams(**kw)


def make(name, spec):
    p = synth.workload_params(spec["workload"], **spec["over"])
    vt = synth.gradient_anomaly_model(p.nz, p.nx, p.dx, p.dz, seed=spec["seed"])
    vs = synth.smooth(vt, spec["smooth"])
    vt32, vs32 = vt.astype(np.float32), vs.astype(np.float32)
    obs = oracle.forward(p, vt32)
    syn, grad = oracle.fwi(p, vs32, obs)
    _, grad64 = oracle.fwi(p, vs32.astype(np.float64), obs.astype(np.float64), dtype=np.float64)
    misfit = oracle.misfit_f32(syn, obs)
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), params=param_vector(p), vel_true=vt32, vel=vs32,
                        obs=obs, syn=syn, grad=grad, grad64=grad64.astype(np.float64), misfit=np.float64(misfit))


if __name__ == "__main__":
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()
    for name, spec in CASES.items():
        make(name, spec)
        print("wrote", name)
