"""CPU oracle for the FWI-2D misfit/gradient path -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it; the product
package ``geopvi_amd`` never does, and has no CPU fallback.

It wraps ``libfwi2d_oracle.so`` (built from ``fwi2d_oracle.c`` by the
Makefile beside it), a C restatement of
geopvi/forward/fwi2d/fdtd2D_modelling.c (finite-difference path), in a float
build that follows the reference's rounding and a double twin.

Parity pins (each function and the check that pins it; see tests/):
  fd_coefficient   exact Fornberg rationals (test_oracle_pins)
  ricker           closed form in double (test_oracle_pins)
  pad_velocity     clamp-index closed form (test_oracle_pins)
  step (stencil)   exact on polynomials of degree <= 2*lc-2 (test_oracle_pins)
  step (ABC)       exact for a translating quadratic wave, Liao order-2
                   transmitting formula (test_oracle_pins)
  forward          first-arrival time in a homogeneous medium, left-right
                   mirror symmetry, f32 vs f64 agreement (test_oracle_physics)
  fwi gradient     zero residual -> zero gradient, linearity in the residual,
                   shot additivity, central finite differences of the misfit
                   (approximate: continuous adjoint) (test_oracle_physics)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GPVI_ORACLE_LIB points the checker at another build of the same source
# (used to run the pin tests against deliberately mutated oracles).
_LIB_PATH = os.environ.get("GPVI_ORACLE_LIB") or os.path.join(_HERE, "libfwi2d_oracle.so")
# The same arithmetic with row-major grid sweeps (identical results, several
# times faster): for long fixture runs.  select(row_major=True) switches.
_LIB_PATH_RM = os.path.join(_HERE, "libfwi2d_oracle_rm.so")
_lib = None
_active = _LIB_PATH


def select(row_major: bool = False) -> str:
    """Use the reference-order build (default) or the row-major build."""
    return use_library(_LIB_PATH_RM if row_major else _LIB_PATH)


def use_library(path: str) -> str:
    """Route every call to the oracle build at ``path`` (another build of
    the same source, e.g. a deliberately mutated one in a mutation test);
    returns the path in use before."""
    global _lib, _active
    prev = _active
    if path != _active:
        _active, _lib = path, None
    return prev


class OParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "nx", "nz", "pml0", "Lc", "laplace_slover", "ns", "nt", "ds", "ns0",
        "depths", "depthr", "nr", "dr", "nr0", "nt_interval")] + [
        (n, ctypes.c_float) for n in ("dx", "dz", "dt", "f0")]


def build() -> str:
    """Compile the oracle library with its Makefile (gcc, OpenMP)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_active):
            raise RuntimeError(f"oracle library missing: {_active} (run `make -C oracle`)")
        L = ctypes.CDLL(_active)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        pp = ctypes.POINTER(OParams)
        L.ofwi_read_params.argtypes = [ctypes.c_char_p, pp]
        L.ofwi_validate.argtypes = [pp]
        L.ofwi_set_threads.argtypes = [ctypes.c_int]
        L.ofwi_fd_coefficient.argtypes = [ctypes.c_int, fp]
        L.ofwi_ricker_f32.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, fp]
        L.ofwi_ricker_f64.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp]
        L.ofwi_pad_velocity_f32.argtypes = [ctypes.c_int] * 3 + [fp, fp]
        L.ofwi_pad_velocity_f64.argtypes = [ctypes.c_int] * 3 + [dp, dp]
        L.ofwi_step_f32.argtypes = [pp, fp, fp, fp, fp]
        L.ofwi_step_f64.argtypes = [pp, dp, dp, dp, dp]
        L.ofwi_fwi_f32.argtypes = [pp, fp, fp, fp, fp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ofwi_fwi_f64.argtypes = [pp, dp, dp, dp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.ofwi_forward_f32.argtypes = [pp, fp, fp]
        L.ofwi_forward_f64.argtypes = [pp, dp, dp]
        _lib = L
    return _lib


def _op(params) -> OParams:
    o = OParams()
    for name, _ in OParams._fields_:
        setattr(o, name, getattr(params, name))
    return o


def _ptr(a: np.ndarray):
    if a.dtype == np.float32:
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    if a.dtype == np.float64:
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    raise TypeError(a.dtype)


def _arr(x, dtype):
    return np.ascontiguousarray(np.asarray(x).reshape(-1), dtype=dtype)


def set_threads(n: int) -> None:
    lib().ofwi_set_threads(int(n))


def get_threads() -> int:
    return int(lib().ofwi_get_threads())


def read_params(path: str) -> OParams:
    o = OParams()
    rc = lib().ofwi_read_params(path.encode(), ctypes.byref(o))
    if rc:
        raise ValueError(f"ofwi_read_params({path}) -> {rc}")
    return o


def validate(params) -> int:
    return int(lib().ofwi_validate(ctypes.byref(_op(params))))


def fd_coefficient(lc: int) -> np.ndarray:
    fd = np.zeros(8, np.float32)
    rc = lib().ofwi_fd_coefficient(int(lc), _ptr(fd))
    if rc:
        raise ValueError(f"unsupported half-width {lc}")
    return fd[:lc].copy()


def ricker(nt: int, dt: float, f0: float, dtype=np.float32) -> np.ndarray:
    out = np.zeros(nt, dtype)
    if dtype == np.float32:
        lib().ofwi_ricker_f32(nt, dt, f0, _ptr(out))
    else:
        lib().ofwi_ricker_f64(nt, float(np.float32(dt)), float(np.float32(f0)), _ptr(out))
    return out


def pad_velocity(params, vel_inner, dtype=np.float32) -> np.ndarray:
    v = _arr(vel_inner, dtype)
    out = np.zeros(params.ntp, dtype)
    fn = lib().ofwi_pad_velocity_f32 if dtype == np.float32 else lib().ofwi_pad_velocity_f64
    fn(params.pml, params.ntx, params.ntz, _ptr(v), _ptr(out))
    return out.reshape(params.ntz, params.ntx)


def step(params, velp, p0, p1, dtype=np.float32) -> np.ndarray:
    """One stencil + ABC step on padded (ntz, ntx) fields; returns p2."""
    v, a, b = (_arr(x, dtype) for x in (velp, p0, p1))
    out = np.zeros(params.ntp, dtype)
    fn = lib().ofwi_step_f32 if dtype == np.float32 else lib().ofwi_step_f64
    fn(ctypes.byref(_op(params)), _ptr(v), _ptr(a), _ptr(b), _ptr(out))
    return out.reshape(params.ntz, params.ntx)


def fwi(params, vel_inner, record_obs, run_fwi: int = 1, dtype=np.float32, shots=None):
    """Misfit-gradient evaluation of one sample: returns (record_syn, grad).

    record_syn/record_obs are flat [ns][nt][nr]; grad is flat [nz][nx]
    (zeroed here, then accumulated over shots as the reference does)."""
    v = _arr(vel_inner, dtype)
    obs = _arr(record_obs, dtype)
    syn = np.zeros(params.ns * params.nt * params.nr, dtype)
    grad = np.zeros(params.nz * params.nx, dtype)
    s0, s1 = (0, -1) if shots is None else shots
    fn = lib().ofwi_fwi_f32 if dtype == np.float32 else lib().ofwi_fwi_f64
    rc = fn(ctypes.byref(_op(params)), _ptr(v), _ptr(syn), _ptr(obs), _ptr(grad), int(run_fwi),
            int(s0), int(s1))
    if rc:
        raise ValueError(f"oracle fwi rejected parameters (code {rc})")
    return syn, grad


def forward(params, vel_inner, dtype=np.float32) -> np.ndarray:
    v = _arr(vel_inner, dtype)
    syn = np.zeros(params.ns * params.nt * params.nr, dtype)
    fn = lib().ofwi_forward_f32 if dtype == np.float32 else lib().ofwi_forward_f64
    rc = fn(ctypes.byref(_op(params)), _ptr(v), _ptr(syn))
    if rc:
        raise ValueError(f"oracle forward rejected parameters (code {rc})")
    return syn


def misfit_f32(syn, obs) -> float:
    """0.5 * sum((syn - obs)**2) with numpy float32 pairwise summation, as
    geopvi/forward/fwi2d/posterior.py:109 computes it."""
    d = np.asarray(syn, np.float32) - np.asarray(obs, np.float32)
    return float(0.5 * np.sum(d ** 2))
