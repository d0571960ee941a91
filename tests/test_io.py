"""Run-file formats: waveform, prior bounds, result files, flow parameters."""
import numpy as np
import pytest
import torch

from geopvi_amd import io, synth
from geopvi_amd.nfvi import FlowsBasedDistribution, Linear, Real2Constr, psvi_mask_2D


@pytest.fixture(autouse=True)
def float64_default():
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def test_waveform_round_trip_and_shape_check(tmp_path):
    p = synth.workload_params("tiny")
    rec = np.arange(p.record_len, dtype=np.float32)
    path = str(tmp_path / "waveform.npy")
    io.save_waveform(path, rec, p)
    arr = np.load(path)
    assert arr.shape == (p.ns, p.nt, p.nr) and arr.dtype == np.float32
    # [ns][nt][nr]: the record of shot s, step t, receiver r
    assert arr[1, 2, 3] == rec[(1 * p.nt + 2) * p.nr + 3]
    assert np.array_equal(io.load_waveform(path, p), rec)
    with pytest.raises(ValueError):
        io.load_waveform(path, synth.workload_params("small"))


def test_prior_bounds_per_depth_row(tmp_path):
    nz, nx, fixed = 5, 3, 2
    b = np.column_stack([1000 + 100 * np.arange(nz), 4000 + 100 * np.arange(nz)])
    path = str(tmp_path / "prior.txt")
    np.savetxt(path, b)
    lo, up = io.load_prior_bounds(path, nz, nx, fixed)
    assert lo.shape == ((nz - fixed) * nx,)
    assert np.array_equal(lo.reshape(nz - fixed, nx)[:, 0], b[fixed:, 0])
    assert np.all(lo.reshape(nz - fixed, nx) == lo.reshape(nz - fixed, nx)[:, :1])
    assert np.array_equal(up.reshape(nz - fixed, nx)[:, 2], b[fixed:, 1])


def test_flow_param_round_trip_and_result_files(tmp_path):
    dim = 20
    mask = psvi_mask_2D(np.ones((3, 3)), dim, nx=5, nz=4)
    f = Linear(dim, "structured", mask=mask)
    with torch.no_grad():
        f.u.normal_()
        f.diag.normal_(0, 0.1)
        f.non_diag.normal_(0, 0.1)
    param = io.get_flow_param(f)
    g = Linear(dim, "structured", mask=mask, param=param)
    assert torch.allclose(g.u, f.u) and torch.allclose(g.diag, f.diag) and torch.allclose(g.non_diag, f.non_diag)
    q = FlowsBasedDistribution([f, Real2Constr(dim, lower=1.0, upper=3.0)],
                               generator=torch.Generator().manual_seed(0))
    paths = io.save_results(str(tmp_path), q, [3.0, 2.0, 1.0], nsamples=50)
    assert np.load(paths["samples"]).shape == (50, dim)
    assert np.allclose(np.load(paths["parameter"]), param)
    assert np.allclose(np.loadtxt(paths["loss"]), [3.0, 2.0, 1.0])
    q2 = FlowsBasedDistribution([Linear(dim, "structured", mask=mask), Real2Constr(dim, lower=1.0, upper=3.0)])
    ck = io.load_model(paths["model"], q2)
    assert ck["iteration"] == 3 and torch.allclose(q2.flows[0].non_diag, f.non_diag)
