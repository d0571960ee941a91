"""Flows of the FWI-2D inversion, device-resident.

Same constructors, parameters and transforms as geopvi/nfvi/flows.py
(Real2Constr :28-77, Constr2Real :80-129, Linear :132-243); bounds are
registered as non-persistent buffers so ``.to(device)`` moves them and
state dicts keep the reference's keys.
"""
from __future__ import annotations

import numpy as np
import torch
import torch.nn as nn


def _bound(value, dim):
    if value is None:
        return None
    if np.isscalar(value):
        value = np.full((dim,), value)
    return torch.as_tensor(np.asarray(value))


class Real2Constr(nn.Module):
    """x in R^dim -> z in (lower, upper) (geopvi/nfvi/flows.py:28-77)."""

    def __init__(self, dim, lower=None, upper=None):
        super().__init__()
        self.dim = dim
        self.register_buffer("lower", _bound(lower, dim), persistent=False)
        self.register_buffer("upper", _bound(upper, dim), persistent=False)

    def forward(self, x, train=True):
        lo, up = self.lower, self.upper
        if lo is None and up is None:
            return x, 0
        if up is None:
            return lo + torch.exp(x), x.sum(axis=-1)
        if lo is None:
            return up - torch.exp(x), x.sum(axis=-1)
        z = lo + (up - lo) / (1 + torch.exp(-x))
        log_det = (torch.log(up - lo) - x - 2 * torch.log(1 + torch.exp(-x))).sum(axis=-1)
        return z, log_det

    def inverse(self, z):
        lo, up = self.lower, self.upper
        if lo is None and up is None:
            return z, 0
        if up is None:
            return torch.log(z - lo), torch.log(1 / (z - lo)).sum(axis=-1)
        if lo is None:
            return torch.log(up - z), torch.log(1 / (up - z)).sum(axis=-1)
        x = torch.log(z - lo) - torch.log(up - z)
        log_det = torch.log(1. / (z - lo) + 1. / (up - z)).sum(axis=-1)
        return x, log_det


class Constr2Real(nn.Module):
    """x in (lower, upper) -> z in R^dim (geopvi/nfvi/flows.py:80-129)."""

    def __init__(self, dim, lower=None, upper=None):
        super().__init__()
        self.dim = dim
        self.register_buffer("lower", _bound(lower, dim), persistent=False)
        self.register_buffer("upper", _bound(upper, dim), persistent=False)

    def forward(self, x, train=True):
        lo, up = self.lower, self.upper
        if lo is None and up is None:
            return x, 0
        if up is None:
            return torch.log(x - lo), torch.log(1 / (x - lo)).sum(axis=-1)
        if lo is None:
            return torch.log(up - x), torch.log(1 / (up - x)).sum(axis=-1)
        z = torch.log(x - lo) - torch.log(up - x)
        log_det = torch.log(1. / (x - lo) + 1. / (up - x)).sum(axis=-1)
        return z, log_det

    def inverse(self, z):
        lo, up = self.lower, self.upper
        if lo is None and up is None:
            return z, 0
        if up is None:
            return lo + torch.exp(z), z.sum(axis=-1)
        if lo is None:
            return up - torch.exp(z), z.sum(axis=-1)
        x = lo + (up - lo) / (1 + torch.exp(-z))
        log_det = (torch.log(up - lo) - z - 2 * torch.log(1 + torch.exp(-z))).sum(axis=-1)
        return x, log_det


def psvi_mask_2D(correlation, ndim, nx=1, nz=1):
    """Band mask of the structured Linear flow for a 2-D correlation
    template (geopvi/utils.py:38-56, examples/fwi2d/fwi2d.py:19-35): one row
    per template entry before the centre, False on the last `offset`
    entries of that row, offset = (cz - iz)*nx + (cx - ix)."""
    z, x = correlation.shape
    rank = int((correlation != 0).sum() // 2)
    cz, cx = (correlation.size // 2) // x, (correlation.size // 2) % x
    mask = np.ones((rank, ndim), dtype=bool)
    i = 0
    for iz in range(z):
        for ix in range(x):
            if correlation[iz, ix] == 0 or iz * x + ix >= correlation.size // 2:
                continue
            off = (cz - iz) * nx + (cx - ix)
            mask[i, ndim - off:] = False
            i += 1
    return mask


class Linear(nn.Module):
    """z = u + L x, L lower triangular with diagonal exp(diag)
    (geopvi/nfvi/flows.py:132-243).

    kernel 'diagonal' (mean-field ADVI), 'structured' (PSVI: band i of L is
    the offset_i-th subdiagonal, L[k, k - o_i] = non_diag[i, k - o_i]) or
    'fullrank'.  mask/param/trainable as in the reference."""

    def __init__(self, dim, kernel="diagonal", mask=None, param=None, trainable=True):
        super().__init__()
        self.dim = dim
        self.kernel = kernel
        self.u = nn.Parameter(torch.zeros(dim))
        self.diag = nn.Parameter(torch.zeros(dim))
        if kernel == "fullrank":
            self.non_diag = nn.Parameter(torch.zeros(int(dim * (dim - 1) / 2)))
        elif kernel == "diagonal":
            self.non_diag = None
        elif kernel == "structured":
            if mask is None:
                mask = np.ones([1, dim])
                mask[0, -1] = 0.
            self.mask = np.asarray(mask)
            self.offset = (self.mask == 0).sum(axis=1)
            self.non_diag = nn.Parameter(torch.zeros(self.mask.shape))
        else:
            raise ValueError(f"unknown kernel {kernel!r}")
        if param is not None:
            param = np.asarray(param)
            if param.size >= dim:
                self.u = nn.Parameter(torch.from_numpy(param[:dim].copy()))
            if param.size >= 2 * dim:
                self.diag = nn.Parameter(torch.from_numpy(np.log(param[dim:2 * dim])))
            if param.size > 2 * dim:
                rest = param[2 * dim:]
                if kernel == "fullrank" and rest.size == int(dim * (dim - 1) / 2):
                    self.non_diag = nn.Parameter(torch.from_numpy(rest.copy()))
                elif kernel == "structured":
                    if rest.size != self.mask.size:
                        raise ValueError("Shape of mask and parameter does not match!")
                    self.non_diag = nn.Parameter(torch.from_numpy(rest.reshape(self.mask.shape).copy()))
        if trainable is False:
            self.u.requires_grad = False
            self.diag.requires_grad = False
            if kernel != "diagonal":
                self.non_diag.requires_grad = False

    def _lower(self, diag):
        L = torch.diag(diag)
        rows, cols = np.tril_indices(self.dim, -1)
        return L.index_put((torch.as_tensor(rows, device=L.device), torch.as_tensor(cols, device=L.device)),
                           self.non_diag, accumulate=True)

    def _bands(self, x):
        """sum_i shift_{o_i}(non_diag[i] * x): band i contributes
        non_diag[i, j] * x[:, j] to row j + o_i."""
        out = torch.zeros_like(x)
        for i, o in enumerate(self.offset):
            o = int(o)
            if o <= 0 or o >= self.dim:
                continue
            out[:, o:] = out[:, o:] + self.non_diag[i, :self.dim - o] * x[:, :self.dim - o]
        return out

    def forward(self, x, train=True):
        diag = torch.exp(self.diag)
        if self.kernel == "fullrank":
            z = self.u + torch.matmul(self._lower(diag), x.T).T
        elif self.kernel == "diagonal":
            z = self.u + diag * x
        else:
            z = self.u + diag * x + self._bands(x)
        log_det = torch.log(diag).sum().repeat(x.shape[0])
        return z, log_det

    def inverse(self, z):
        diag = torch.exp(self.diag)
        b = z - self.u
        if self.kernel == "diagonal":
            x = b / diag
        elif self.kernel == "fullrank":
            x = torch.linalg.solve_triangular(self._lower(diag), b.T, upper=False).T
        else:
            # Banded lower-triangular solve on the host (not on the training
            # path; the reference also solves with scipy, flows.py:236-240).
            import scipy.sparse as sparse
            import scipy.sparse.linalg as linalg
            offs = [0] + [-int(o) for o in self.offset]
            diags = [diag.detach().cpu().numpy()] + [self.non_diag[i].detach().cpu().numpy()
                                                      for i in range(len(self.offset))]
            L = sparse.diags(diags, offs, shape=(self.dim, self.dim), format="csr")
            sol = linalg.spsolve_triangular(L, b.detach().cpu().numpy().T, lower=True)
            x = torch.as_tensor(np.ascontiguousarray(sol.T), dtype=b.dtype, device=b.device)
        log_det = -torch.log(diag).sum().repeat(x.shape[0])
        return x, log_det
