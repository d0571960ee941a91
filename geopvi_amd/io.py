"""Files of a 2-D FWI inversion run, in the reference's formats.

* observed data ``waveform.npy``: float32 records, shape (ns, nt, nr)
  (fdtd2D_modelling.c:273 record layout; read flattened by
  examples/fwi2d/fwi2d.py:128-129);
* prior bounds ``prior.txt``: one (lower, upper) row per depth row, the
  fixed top layer included (examples/fwi2d/fwi2d.py:131-136);
* results (examples/fwi2d/fwi2d.py:208-229): ``{flow}_{kernel}_samples.npy``
  (posterior samples), ``{flow}_{kernel}_parameter.npy`` (the Linear flow's
  [u, exp(diag), non_diag] vector, the layout ``Linear(param=...)``
  reads back), ``{flow}_{kernel}_loss.txt`` and ``{flow}_{kernel}_model.pt``
  ({'iteration', 'model_state_dict', 'loss'}).
"""
from __future__ import annotations

import os
from typing import Optional, Tuple

import numpy as np

from .forward.fwi2d.params import FWIParams


def save_waveform(path: str, records, params: FWIParams) -> None:
    """Records of one model (ns*nt*nr values, [ns][nt][nr] order) -> .npy (ns, nt, nr) float32."""
    r = np.asarray(records, dtype=np.float32).reshape(params.ns, params.nt, params.nr)
    np.save(path, r)


def load_waveform(path: str, params: Optional[FWIParams] = None) -> np.ndarray:
    """.npy records -> flat float32 (ns*nt*nr,); shape checked against params if given."""
    r = np.load(path, allow_pickle=False)
    if params is not None and r.size != params.record_len:
        raise ValueError(f"{path}: {r.shape} holds {r.size} values, expected (ns, nt, nr) = "
                         f"({params.ns}, {params.nt}, {params.nr})")
    return np.ascontiguousarray(r, dtype=np.float32).reshape(-1)


def load_prior_bounds(path: str, nz: int, nx: int, layer_fixed: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """prior.txt (nz rows of lower, upper) -> per-cell (lower, upper) of the
    free cells (rows layer_fixed..nz-1, x fastest), float64."""
    b = np.loadtxt(path, ndmin=2)
    if b.shape[0] != nz or b.shape[1] < 2:
        raise ValueError(f"{path}: expected {nz} rows of (lower, upper), got {b.shape}")
    lower = np.broadcast_to(b[layer_fixed:, 0].astype(np.float64)[:, None], (nz - layer_fixed, nx)).reshape(-1)
    upper = np.broadcast_to(b[layer_fixed:, 1].astype(np.float64)[:, None], (nz - layer_fixed, nx)).reshape(-1)
    return lower.copy(), upper.copy()


def get_flow_param(flow) -> np.ndarray:
    """[u, exp(diag)(, non_diag flattened)] of a Linear flow."""
    mus = flow.u.detach().cpu().numpy()
    sigmas = np.exp(flow.diag.detach().cpu().numpy())
    if flow.non_diag is None:
        return np.hstack([mus, sigmas])
    return np.hstack([mus, sigmas, flow.non_diag.detach().cpu().numpy().reshape(-1)])


def result_paths(outdir: str, flow: str = "Linear", kernel: str = "structured") -> dict:
    stem = os.path.join(outdir, f"{flow}_{kernel}")
    return {"samples": stem + "_samples.npy", "parameter": stem + "_parameter.npy", "loss": stem + "_loss.txt",
            "model": stem + "_model.pt"}


def save_results(outdir: str, variational, loss, flow: str = "Linear", kernel: str = "structured",
                 nsamples: int = 2000, linear_index: int = -2) -> dict:
    """Write the run's four result files; returns their paths.  The Linear
    flow whose parameters are saved is variational.flows[linear_index]
    (the one before the final Real2Constr)."""
    import torch

    os.makedirs(outdir, exist_ok=True)
    paths = result_paths(outdir, flow, kernel)
    variational.eval()
    np.save(paths["samples"], variational.sample(nsamples))
    np.save(paths["parameter"], get_flow_param(variational.flows[linear_index]))
    np.savetxt(paths["loss"], np.asarray(loss, dtype=np.float64))
    torch.save({"iteration": len(loss), "model_state_dict": variational.state_dict(),
                "loss": [float(v) for v in loss]}, paths["model"])
    return paths


def load_model(path: str, variational) -> dict:
    """Restore a model written by save_results (tensors and plain values only)."""
    import torch

    ck = torch.load(path, map_location="cpu", weights_only=True)
    variational.load_state_dict(ck["model_state_dict"])
    return ck
