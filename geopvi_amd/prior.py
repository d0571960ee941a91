"""Prior of the FWI-2D inversion (geopvi/prior.py Uniform :63-92) and the
smoothing operator it may carry (geopvi/utils.py:5-22), on the device."""
from __future__ import annotations

import numpy as np
import scipy.sparse as sparse
import torch


def delta(n):
    """Second-difference matrix with one-sided ends (geopvi/utils.py:5-10)."""
    d0 = np.full((n,), -2.0)
    d0[0] = -1.0
    d0[-1] = -1.0
    d1 = np.full((n - 1,), 1.0)
    return sparse.diags([d0, d1, d1], [0, -1, 1]).tocsc()


def smooth_matrix_2D(nx, nz, smoothx, smoothz):
    """[I_z/sx (x) D_x ; D_z/sz (x) I_x] (geopvi/utils.py:12-22)."""
    sx = np.full((nz,), smoothx, dtype=float)
    sz = np.full((nz,), smoothz, dtype=float)
    dz = delta(nz) / sz[:, None]
    Sx = sparse.kron(sparse.eye(nz) / sx, delta(nx))
    Sz = sparse.kron(dz, sparse.eye(nx))
    return sparse.vstack([Sx, Sz])


def _to_torch_sparse(S, device, dtype):
    S = S.tocoo()
    idx = torch.as_tensor(np.vstack([S.row, S.col]), dtype=torch.long)
    return torch.sparse_coo_tensor(idx, torch.as_tensor(S.data, dtype=dtype), S.shape).coalesce().to(device)


class Uniform:
    """Uniform prior on [lower, upper] (geopvi/prior.py:63-92):
    log p(x) = -sum(log(upper - lower)) [- 0.5 ||S x||^2 with a smoothing
    matrix S].  Works on CPU or CUDA inputs; the smoothing term and its
    gradient are computed on the input's device."""

    def __init__(self, lower=np.array([0.]), upper=np.array([1.]), smooth_matrix=None):
        if np.isscalar(upper):
            upper = np.array([upper])
        if np.isscalar(lower):
            lower = np.array([lower])
        self.lower = torch.from_numpy(np.asarray(lower))
        self.upper = torch.from_numpy(np.asarray(upper))
        self.smooth_matrix = smooth_matrix
        self._const = {}
        self._S = {}

    def _logz(self, device, dtype):
        key = (str(device), dtype)
        if key not in self._const:
            self._const[key] = -torch.log(self.upper - self.lower).sum(axis=-1).to(device=device, dtype=dtype)
        return self._const[key]

    def log_prob(self, x):
        logp = self._logz(x.device, x.dtype)
        if self.smooth_matrix is None:
            return logp
        key = (str(x.device), x.dtype)
        if key not in self._S:
            self._S[key] = _to_torch_sparse(self.smooth_matrix, x.device, x.dtype)
        t = torch.sparse.mm(self._S[key], x.T)
        return logp - 0.5 * (t ** 2).sum(axis=0)
