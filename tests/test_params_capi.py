"""Host-side checks that need no GPU: the parameter-file format, validation
error codes, and that libgeopvi_fwi2d.so loads and exports every function
include/geopvi_fwi2d.h declares."""
import configparser
import ctypes
import os
import re

import numpy as np
import pytest

from geopvi_amd import _lib, synth
from geopvi_amd.forward.fwi2d.params import CParams, FWIParams, prepare_fwi_parameters
from geopvi_amd.forward.fwi2d import solver

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "geopvi_fwi2d.h")

# The [FWI] section of the reference's example configuration
# (examples/fwi2d/input/config.ini), as a configparser would read it.
EXAMPLE_FWI = dict(nx="250", nz="110", layer_fixed="10", vel_fixed="1950", pml="10", Lc="3", laplace_slover="0",
                   ns="12", nt="2001", ds="20", ns0="15", depths="1", depthr="9", nr="250", dr="1", nr0="0",
                   nt_interval="2", dx="20.0", dz="20.0", dt="0.002", f0="10.0")


def example_config():
    cfg = configparser.ConfigParser()
    cfg["FWI"] = EXAMPLE_FWI
    return cfg


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|size_t|char)\s*\**\s*(\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_function():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 25, names
    for n in names:
        assert hasattr(lib, n), n
    # and the Python binding binds exactly those
    assert sorted(_lib.SIGNATURES) == names


def test_prepare_fwi_parameters_round_trip(tmp_path, oracle_mod):
    path = str(tmp_path / "input_params.txt")
    prepare_fwi_parameters(path, example_config())
    lines = open(path).read().splitlines()
    assert len(lines) == 38 and lines[0].startswith("--") and lines[1] == "250"
    p = FWIParams.read(path)
    assert p == FWIParams.from_config(example_config())
    # the C parser of the library and the oracle's read the same values
    c = CParams()
    assert _lib.load().gpvi_fwi2d_read_params(path.encode(), ctypes.byref(c)) == 0
    assert FWIParams.from_c(c) == p
    o = oracle_mod.read_params(path)
    for name, _ in o._fields_:
        assert getattr(o, name) == pytest.approx(getattr(p, name)), name


def test_param_file_blank_lines_and_errors(tmp_path):
    p = synth.workload_params("c0")
    text = p.to_text().replace("\n--Total time", "\n\n--Total time")  # blank line before a comment
    path = tmp_path / "p.txt"
    path.write_text(text)
    assert FWIParams.read(str(path)) == p
    c = CParams()
    assert _lib.load().gpvi_fwi2d_read_params(str(path).encode(), ctypes.byref(c)) == 0
    assert FWIParams.from_c(c) == p
    rc = _lib.load().gpvi_fwi2d_read_params(str(tmp_path / "missing.txt").encode(), ctypes.byref(c))
    assert rc == -4 and "cannot open" in _lib.last_error()
    (tmp_path / "short.txt").write_text("\n".join(p.to_text().splitlines()[:10]) + "\n")
    assert _lib.load().gpvi_fwi2d_read_params(str(tmp_path / "short.txt").encode(), ctypes.byref(c)) == -4


@pytest.mark.parametrize("change,code", [
    (dict(laplace_slover=2), -2),
    (dict(Lc=7), -2),
    (dict(Lc=2), -2),
    (dict(ns0=500), -3),
    (dict(depthr=-100), -3),
    (dict(nr0=-30), -3),
    (dict(dr=0), -1),
    (dict(nt=0), -1),
    (dict(dt=0.0), -1),
    (dict(nx=24000, nz=24000), -1),        # a padded field of 1 GiB or more
    (dict(nx=16200, nz=16500), -1),        # just over 1 GiB
    (dict(dx=float("inf")), -1),
    (dict(dt=float("nan")), -1),
    (dict(nt=1 << 25, nr=64), -1),         # 2^31 record values per shot
])
def test_validation_codes(change, code):
    if change.get("nx") == 16200:
        ok = synth.workload_params("c0", nx=15000, nz=16000).to_c()   # 15028 x 16028 padded: under 1 GiB
        assert _lib.load().gpvi_fwi2d_validate(ctypes.byref(ok)) == 0
    p = synth.workload_params("c0", **change)
    c = p.to_c()
    assert _lib.load().gpvi_fwi2d_validate(ctypes.byref(c)) == code
    assert _lib.last_error()
    with pytest.raises(_lib.GpviError):
        solver.validate(p)


def test_pseudo_spectral_parameters_accepted():
    # laplace_slover = 1 drops the stencil half-width from the padding
    # (pml = pml0), so Lc is not checked and a source may sit closer to the edge
    p = synth.workload_params("c0", laplace_slover=1, Lc=9)
    assert _lib.load().gpvi_fwi2d_validate(ctypes.byref(p.to_c())) == 0
    q = synth.workload_params("c0", laplace_slover=1, ns0=-p.pml0)
    assert _lib.load().gpvi_fwi2d_validate(ctypes.byref(q.to_c())) == 0
    q = synth.workload_params("c0", laplace_slover=1, ns0=-p.pml0 - 1)
    assert _lib.load().gpvi_fwi2d_validate(ctypes.byref(q.to_c())) == -3


def test_workspace_sizes():
    p = synth.workload_params("c1")
    full = solver.batch_workspace_bytes(p, 16, 0)
    ck = solver.batch_workspace_bytes(p, 16, 100)
    # full store: one inner field per nt_interval steps per unit
    assert full > 16 * 20 * p.nsave * p.ni * 4
    assert ck < full / 5
    assert solver.plain_workspace_bytes(p) > p.nsave * p.ni * 4
    # forward modelling only: four rotating fields per unit, no store
    model = solver.model_workspace_bytes(p, 16)
    fe = (p.nz + 2 * (p.pml0 + p.Lc + 1) + 32) * 448
    assert 16 * 20 * 4 * fe * 4 <= model < 16 * 20 * 5 * fe * 4 + 16 * fe * 4 + (1 << 24)


def test_tuning_round_trip():
    v, g, s = _lib.get_tuning()
    try:
        _lib.set_tuning(2, 1024, 16)
        assert _lib.get_tuning() == (2, 1024, 16)
        assert _lib.tile_shape(2) == (64, 64)
        with pytest.raises(_lib.GpviError):
            _lib.set_tuning(99, 0, 16)
    finally:
        _lib.set_tuning(v, g, s)


def test_lanes_round_trip():
    lanes = _lib.get_lanes()
    assert lanes == 2   # the default: two streams per batched call
    try:
        for n in (1, 2, 3, 4):
            _lib.set_lanes(n)
            assert _lib.get_lanes() == n
        for bad in (0, 5, -1):
            with pytest.raises(_lib.GpviError):
                _lib.set_lanes(bad)
        assert _lib.get_lanes() == 4
    finally:
        _lib.set_lanes(lanes)


def test_compat_entry_reports_bad_parameter_file(tmp_path):
    lib = _lib.load()
    buf = (ctypes.c_float * 4)()
    lib.fwi_2D(str(tmp_path / "nope.txt").encode(), buf, buf, buf, buf, 1, 0)
    assert lib.gpvi_compat_last_status() == -4


def test_pseudo_spectral_batch_unit_limit():
    # the pseudo-spectral kernels launch one grid row per (sample, shot)
    # unit: more than 65535 units is refused before any device work
    p = synth.workload_params("tiny", laplace_slover=1)
    c = p.to_c()
    fake = ctypes.c_void_p(256)  # never dereferenced: the call returns first
    rc = _lib.load().gpvi_fwi2d_batch(ctypes.byref(c), 40000, fake, fake, fake, fake, None, 1, 0, 0, 0, fake,
                                      1 << 40, None)
    assert rc == -1 and "65535" in _lib.last_error()
