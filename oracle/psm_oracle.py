"""CPU oracle of the pseudo-spectral (PSM) Laplacian path -- TEST INFRASTRUCTURE ONLY.

numpy restatement of geopvi/forward/fwi2d/fdtd2D_modelling.c with
laplace_slover = 1: Lc = 0, so pml = pml0 (fdtd.c:75-85); every step
replaces the finite-difference stencil by pstd_2d_calculate_p
(fdtd.c:783-830) -- a double-precision 2-D DFT of p1, multiplication by
-|k|^2 (wavenumber, fdtd.c:1127-1184), the unnormalised inverse DFT, and

    p2 = (float)((double)(2*p1 - p0) + (double)powf(v*dt, 2) * Re / (ntx*ntz))

-- followed by the same ordered absorbing sweeps (over [0, pml0) bands),
source/residual injection, wavefield store and imaging as the FD path.
FFTW's conventions: FFTW_FORWARD is the unscaled e^{-i...} transform,
FFTW_BACKWARD the unscaled e^{+i...} one.  numpy's FFT (pocketfft) computes
the same transforms in double; the two libraries agree to rounding.

Pins (tests/test_psm_oracle.py): the Laplacian of a sampled plane wave
cos(2 pi m x / L) is -(2 pi m / L)^2 times it (exact for a band-limited
field), the wavenumber table against its closed form, left-right mirror
symmetry of the records, zero residual -> zero gradient, linearity of the
gradient in the residual, and closeness to the high-order FD oracle for a
well-resolved wave.
"""
from __future__ import annotations

import numpy as np

import oracle as _fd

PI = 3.1415926  # the reference's pi (fdtd2D_modelling.c:15)
F32 = np.float32


def geometry(p):
    pml = p.pml0
    ntx, ntz = p.nx + 2 * pml, p.nz + 2 * pml
    s_ix = np.array([is_ * p.ds + pml + p.ns0 for is_ in range(p.ns)])
    r_ix = np.array([pml + ir * p.dr + p.nr0 for ir in range(p.nr)])
    return pml, ntx, ntz, s_ix, pml + p.depths, r_ix, pml + p.depthr


def wavenumber(ntx, ntz, dx, dz):
    """|k|^2 per point, float, [ntz, ntx] (fdtd.c:1127-1184)."""
    dkz = F32(1.0 / ntz / float(F32(dz)))
    dkx = F32(1.0 / ntx / float(F32(dx)))
    az = ntz // 2 + (ntz % 2)
    ax = ntx // 2 + (ntx % 2)
    iz = np.arange(ntz)
    ix = np.arange(ntx)
    kz = np.where(iz < az, 2 * PI * float(dkz) * iz, 2 * PI * float(dkz) * (ntz - iz)).astype(F32)
    kx = np.where(ix < ax, 2 * PI * float(dkx) * ix, 2 * PI * float(dkx) * (ntx - ix)).astype(F32)
    k = kx[None, :].astype(np.float64) ** 2 + kz[:, None].astype(np.float64) ** 2
    return k.astype(F32)


def laplacian_unnormalised(p1, k):
    """Re(IDFT(-k * DFT(p1))) without the 1/(ntx*ntz) (double)."""
    F = np.fft.fft2(p1.astype(np.float64))
    F = -k.astype(np.float64) * F
    return np.fft.ifft2(F, norm="forward").real


class _Ctx:
    def __init__(self, p, vel_inner):
        self.p = p
        self.pml, self.ntx, self.ntz, self.s_ix, self.s_iz, self.r_ix, self.r_iz = geometry(p)
        self.dt, self.dx = F32(p.dt), F32(p.dx)
        v = np.asarray(vel_inner, F32).reshape(p.nz, p.nx)
        self.v = np.pad(v, self.pml, mode="edge")
        self.k = wavenumber(self.ntx, self.ntz, p.dx, p.dz)
        self.N = float(self.ntx * self.ntz)
        self.vdt2 = (self.v * self.dt) * (self.v * self.dt)
        s = F32(1.0) * self.v * self.dt / self.dx
        self.t11 = (F32(2) - s) * (F32(1) - s) / F32(2)
        self.t12 = s * (F32(2) - s)
        self.t13 = s * (s - F32(1)) / F32(2)
        self.w = np.array([1.0 - 1.0 * i / p.pml0 for i in range(p.pml0)]).astype(F32)
        self.rik = _fd.ricker(p.nt, p.dt, p.f0)

    def _h(self, p0, p1, sl, d, axis):
        """Transmitting extrapolation for the band points `sl`, inward step d."""
        def sh(a, m):
            return np.roll(a, -m * d, axis=axis)[sl]
        t11, t12, t13 = self.t11[sl], self.t12[sl], self.t13[sl]
        return (F32(2) * t11 * p1[sl] + F32(2) * t12 * sh(p1, 1) + F32(2) * t13 * sh(p1, 2)
                - t11 * t11 * p0[sl] - F32(2) * t11 * t12 * sh(p0, 1)
                - (t12 * t12 + F32(2) * t11 * t13) * sh(p0, 2)
                - F32(2) * t12 * t13 * sh(p0, 3) - t13 * t13 * sh(p0, 4))

    @staticmethod
    def _blend(wt, h, u):
        return ((wt * h).astype(np.float64) + (1.0 - wt.astype(np.float64)) * u.astype(np.float64)).astype(F32)

    def step(self, p0, p1):
        """PSM update then the four ordered absorbing sweeps."""
        lap = laplacian_unnormalised(p1, self.k)
        base = (F32(2) * p1 - p0).astype(np.float64)
        p2 = (base + self.vdt2.astype(np.float64) * lap / self.N).astype(F32)
        pml, ntx, ntz = self.pml, self.ntx, self.ntz
        rows = np.arange(pml)
        # top (rows 0..pml-1, inward +z), bottom (rows ntz-pml.., inward -z)
        sl = (slice(0, pml), slice(None))
        p2[sl] = self._blend(self.w[rows][:, None], self._h(p0, p1, sl, 1, 0), p2[sl])
        sl = (slice(ntz - pml, ntz), slice(None))
        wb = self.w[ntz - np.arange(ntz - pml, ntz) - 1][:, None]
        p2[sl] = self._blend(wb, self._h(p0, p1, sl, -1, 0), p2[sl])
        # left (cols 0..pml-1, inward +x), right (inward -x)
        sl = (slice(None), slice(0, pml))
        p2[sl] = self._blend(self.w[rows][None, :], self._h(p0, p1, sl, 1, 1), p2[sl])
        sl = (slice(None), slice(ntx - pml, ntx))
        wr = self.w[ntx - np.arange(ntx - pml, ntx) - 1][None, :]
        p2[sl] = self._blend(wr, self._h(p0, p1, sl, -1, 1), p2[sl])
        return p2


def _forward_shot(c, is_, run_fwi, keep_store):
    p, dt = c.p, c.dt
    p0 = np.zeros((c.ntz, c.ntx), F32)
    p1 = np.zeros_like(p0)
    rec = np.zeros((p.nt, p.nr), F32)
    store = []
    for it in range(p.nt):
        p2 = c.step(p0, p1)
        s = (c.s_iz, c.s_ix[is_])
        p2[s] = p2[s] + c.rik[it] * dt * dt * c.v[s] * c.v[s]
        if keep_store and it % p.nt_interval == 0:
            store.append(((p2 - F32(2) * p1) + p0) / dt / dt if run_fwi else p2.copy())
        rec[it] = p2[c.r_iz, c.r_ix]
        p0, p1 = p1, p2
    return rec, store


def fwi(p, vel_inner, record_obs, run_fwi=1):
    """(record_syn [ns][nt][nr], grad [nz][nx] float32), like fwi_2D with laplace_slover = 1."""
    c = _Ctx(p, vel_inner)
    obs = np.asarray(record_obs, F32).reshape(p.ns, p.nt, p.nr)
    syn = np.zeros_like(obs)
    grad = np.zeros((p.nz, p.nx), F32)
    dt, k = c.dt, p.nt_interval
    pml = c.pml
    inner = (slice(pml, c.ntz - pml), slice(pml, c.ntx - pml))
    v3 = (c.v.astype(np.float64) ** 3).astype(F32)
    for is_ in range(p.ns):
        rec, store = _forward_shot(c, is_, run_fwi, True)
        syn[is_] = rec
        res = (rec - obs[is_]) if run_fwi else obs[is_]
        p0 = np.zeros((c.ntz, c.ntx), F32)
        p1 = np.zeros_like(p0)
        for it in range(p.nt - 1, -1, -1):
            p2 = c.step(p0, p1)
            for ir in range(p.nr):
                r = (c.r_iz, c.r_ix[ir])
                p2[r] = p2[r] + res[it, ir] * dt * dt * c.v[r] * c.v[r]
            if it % k == 0:
                src = store[it // k][inner]
                u = p2[inner]
                if run_fwi:
                    term = (-2.0 / v3[inner].astype(np.float64) * src.astype(np.float64)
                            * u.astype(np.float64) * k * float(dt))
                else:
                    term = (src * u * F32(k) * dt).astype(np.float64)
                grad = (grad.astype(np.float64) + term).astype(F32)
            p0, p1 = p1, p2
    return syn.reshape(-1), grad.reshape(-1)


def forward(p, vel_inner):
    """record_syn of every shot (forward_2D with laplace_slover = 1)."""
    c = _Ctx(p, vel_inner)
    return np.concatenate([_forward_shot(c, is_, 1, False)[0].reshape(-1) for is_ in range(p.ns)])
