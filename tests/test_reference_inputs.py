"""The input files the reference ships with its 2-D FWI example, read the way
its driver reads them (examples/fwi2d/fwi2d.py:106-168), copied under
tests/golden/reference_inputs/ by tests/golden/copy_reference_inputs.py.

* config.ini through configparser (extended interpolation, as the driver
  sets it) and prepare_fwi_parameters; the positional file it writes is read
  back by the Python parser, the C ABI's gpvi_fwi2d_read_params and the
  oracle's parser, which must agree with each other and with the INI values.
* prior.txt through io.load_prior_bounds with the example's fixed water
  layer (10 rows): the driver's lower/upper vectors (fwi2d.py:131-136).
* Linear_diagonal_parameter.npy (a saved mean-field ADVI posterior of the
  example, [u, exp(diag)] over 25 000 cells; fwi2d.py:40-48, 217-219) loaded
  without pickle into Linear(param=...) and mapped through Real2Constr with
  the prior bounds: a velocity model inside the bounds, which the solver's
  validation and stability limit accept at the example's dx and dt.
"""
import configparser
import ctypes
import os

import numpy as np
import pytest
import torch

from geopvi_amd import io
from geopvi_amd.forward.fwi2d.params import FWIParams, prepare_fwi_parameters
from geopvi_amd.nfvi import Linear, Real2Constr

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "reference_inputs")


def example_config():
    cfg = configparser.ConfigParser()
    cfg._interpolation = configparser.ExtendedInterpolation()
    assert cfg.read(os.path.join(INPUTS, "config.ini"))
    return cfg


def example_model():
    """(params, full velocity model (nz, nx) float64, mask of free cells,
    per-cell posterior std in logit space): the saved posterior's mean
    through Real2Constr, the water layer at vel_fixed."""
    cfg = example_config()
    p = FWIParams.from_config(cfg)
    layer, vel_fixed = cfg.getint("FWI", "layer_fixed"), cfg.getfloat("FWI", "vel_fixed")
    lo, up = io.load_prior_bounds(os.path.join(INPUTS, "prior.txt"), p.nz, p.nx, layer)
    ndim = (p.nz - layer) * p.nx
    param = np.load(os.path.join(INPUTS, "Linear_diagonal_parameter.npy"), allow_pickle=False)
    lin = Linear(ndim, kernel="diagonal", param=param)
    with torch.no_grad():
        z, _ = Real2Constr(ndim, lower=lo, upper=up).forward(lin.u.detach().reshape(1, -1))
    v = np.full((p.nz, p.nx), vel_fixed)
    v[layer:] = z.numpy().reshape(p.nz - layer, p.nx)
    mask = np.ones((p.nz, p.nx), bool)
    mask[:layer] = False
    return p, v, mask, np.exp(lin.diag.detach().numpy())


def test_config_through_every_parser(tmp_path, oracle_mod, gpvi_lib):
    cfg = example_config()
    path = str(tmp_path / "input_params.txt")
    prepare_fwi_parameters(path, cfg)
    py = FWIParams.read(path)
    assert py == FWIParams.from_config(cfg)
    c = FWIParams(nx=1, nz=1).to_c()
    assert gpvi_lib.gpvi_fwi2d_read_params(path.encode(), ctypes.byref(c)) == 0
    assert FWIParams.from_c(c) == py
    o = oracle_mod.read_params(path)
    assert all(getattr(o, f) == getattr(py, f) for f, _ in o._fields_)
    # the example's values (examples/fwi2d/input/config.ini)
    assert (py.nx, py.nz, py.ns, py.nt, py.nr, py.ds, py.ns0, py.depths, py.depthr) == \
        (250, 110, 12, 2001, 250, 20, 15, 1, 9)
    assert (py.dx, py.dz, py.dt, py.f0, py.nt_interval, py.pml0, py.Lc) == \
        (20.0, 20.0, np.float32(0.002), 10.0, 2, 10, 3)
    assert gpvi_lib.gpvi_fwi2d_validate(ctypes.byref(py.to_c())) == 0


def test_prior_bounds_as_the_driver_builds_them():
    cfg = example_config()
    nz, nx, layer = cfg.getint("FWI", "nz"), cfg.getint("FWI", "nx"), cfg.getint("FWI", "layer_fixed")
    lo, up = io.load_prior_bounds(os.path.join(INPUTS, "prior.txt"), nz, nx, layer)
    rows = np.loadtxt(os.path.join(INPUTS, "prior.txt"))
    assert rows.shape == (110, 2) and lo.shape == up.shape == ((nz - layer) * nx,)
    # cell (iz, ix) of the free block takes row layer + iz's bounds
    for iz, ix in ((0, 0), (0, nx - 1), (37, 120), (nz - layer - 1, 5)):
        k = iz * nx + ix
        assert lo[k] == rows[layer + iz, 0] and up[k] == rows[layer + iz, 1]
    assert np.all(lo < up)


def test_saved_posterior_maps_to_a_valid_model():
    p, v, mask, sigma = example_model()
    cfg = example_config()
    lo, up = io.load_prior_bounds(os.path.join(INPUTS, "prior.txt"), p.nz, p.nx, cfg.getint("FWI", "layer_fixed"))
    free = v[mask]
    assert free.shape == (25000,) and np.all(free > lo) and np.all(free < up)
    assert np.all(sigma > 0)
    # deeper rows are faster on average (the example's velocity trend)
    rowmean = v[10:].mean(axis=1)
    assert rowmean[-10:].mean() > rowmean[:10].mean() + 500
    assert v.max() < p.max_stable_velocity()
    # the saved vector round-trips through the flow (io.get_flow_param layout)
    param = np.load(os.path.join(INPUTS, "Linear_diagonal_parameter.npy"), allow_pickle=False)
    lin = Linear(25000, kernel="diagonal", param=param)
    np.testing.assert_allclose(io.get_flow_param(lin), param, rtol=1e-14)
