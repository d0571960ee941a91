"""CPU oracle of the variational caller chain -- TEST INFRASTRUCTURE ONLY.

numpy (float64) restatements of the transforms the 2-D FWI inversion
composes, each citing the reference lines it follows:

  linear_forward     geopvi/nfvi/flows.py:205-220 (diagonal / structured /
                     fullrank z = u + L x, log|det| = sum(diag))
  real2constr        geopvi/nfvi/flows.py:48-62 (logistic map to (lo, up))
  constr2real        geopvi/nfvi/flows.py:100-113
  uniform_log_prob   geopvi/prior.py:83-92
  log_prob_base      geopvi/nfvi/models.py:29-33

Pins (tests/test_nfvi.py): the structured kernel against an explicitly
built dense lower-triangular L, log|det| against numpy's slogdet of that L,
forward/inverse round trips, the logistic Jacobian against central finite
differences, psvi_mask_2D offsets against the template geometry, and the
smoothing matrix against constant and linear fields (zero second
differences).
"""
from __future__ import annotations

import numpy as np


def structured_dense(diag_log, non_diag, offsets):
    """Dense L of a structured Linear flow: exp(diag) on the diagonal and
    L[k, k - o_i] = non_diag[i, k - o_i] (band i, offset o_i)."""
    dim = diag_log.size
    L = np.diag(np.exp(diag_log))
    for i, o in enumerate(offsets):
        for k in range(o, dim):
            L[k, k - o] += non_diag[i, k - o]
    return L


def linear_forward(u, diag_log, x, kernel="diagonal", non_diag=None, offsets=None):
    diag = np.exp(diag_log)
    if kernel == "diagonal":
        z = u + diag * x
    elif kernel == "structured":
        z = u + diag * x
        for i, o in enumerate(offsets):
            z[:, o:] += non_diag[i, :u.size - o] * x[:, :u.size - o]
    elif kernel == "fullrank":
        dim = u.size
        L = np.diag(diag)
        L[np.tril_indices(dim, -1)] = non_diag
        z = u + (L @ x.T).T
    else:
        raise ValueError(kernel)
    return z, np.full(x.shape[0], np.log(diag).sum())


def real2constr(lower, upper, x):
    z = lower + (upper - lower) / (1 + np.exp(-x))
    log_det = (np.log(upper - lower) - x - 2 * np.log(1 + np.exp(-x))).sum(axis=-1)
    return z, log_det


def constr2real(lower, upper, z):
    x = np.log(z - lower) - np.log(upper - z)
    log_det = np.log(1. / (z - lower) + 1. / (upper - z)).sum(axis=-1)
    return x, log_det


def uniform_log_prob(lower, upper, x, S=None):
    logp = -np.log(upper - lower).sum()
    if S is None:
        return np.full(x.shape[0], logp)
    t = S @ x.T
    return logp - 0.5 * (np.asarray(t) ** 2).sum(axis=0)


def log_prob_base(x, base="Normal"):
    if base == "Normal":
        return -0.5 * x.shape[1] * np.log(2 * np.pi) - 0.5 * (x ** 2).sum(axis=1)
    return np.zeros(x.shape[0])
