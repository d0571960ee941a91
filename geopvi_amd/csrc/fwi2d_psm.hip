// Pseudo-spectral Laplacian path (laplace_slover = 1) of the FWI-2D solver
// for gfx950: geopvi/forward/fwi2d/fdtd2D_modelling.c with Lc = 0
// (pml = pml0, fdtd.c:75-85) and pstd_2d_calculate_p (fdtd.c:783-830) in
// place of the stencil.  One time step of every unit of a batch:
//
//   pack    p1 -> double field                    (fdtd.c:791-799); only at
//           the first step of a pass: within a pass the point kernel of
//           the previous step leaves its result there
//   FFT     batched 2-D forward transform         (hipFFT, D2Z)
//   scale   multiply by -|k|^2                    (fdtd.c:805-813)
//   IFFT    batched 2-D inverse (unnormalised)    (hipFFT, Z2D)
//   point   p2 = 2 p1 - p0 + (v dt)^2 Re / (ntx ntz), then the four ordered
//           absorbing sweeps over the pml0-wide bands, source or residual
//           injection, store/records or imaging   (fdtd.c:819-826, 833-978)
//
// The reference transforms the real field as a complex one (fftw c2c) and
// keeps the real part of the inverse.  The spectrum of a real field is
// Hermitian and |k|^2 is even in both axes (kx(ix) = kx(ntx - ix)), so the
// scaled spectrum stays Hermitian and its inverse is real: the
// real-to-complex / complex-to-real pair computes the same values with half
// the transform work; only the rounding differs, as between any two FFT
// libraries.  Transforms are in place, double precision (FFTW's
// fftw_complex is double); real rows are padded to rp = 2 (ntx/2 + 1)
// doubles, the spectrum has ntx/2 + 1 complex columns.  Every other
// operation keeps the reference's float/double typing.
#include "fwi2d_kernels.h"
#include "fwi2d_step.h"

namespace gpvi {

constexpr double kPi = 3.1415926;  // the reference's pi (fdtd2D_modelling.c:15)

// |k|^2 per padded point (wavenumber, fdtd.c:1127-1184): float dkx/dkz from
// double expressions, float kx/kz, double squares summed and rounded.
__global__ void psm_wavenumber_kernel(float* __restrict__ k, int ntx, int ntz, float dx, float dz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntx * ntz) return;
    const int iz = i / ntx, ix = i - iz * ntx;
    const float dkz = (float)(1.0 / ntz / (double)dz);
    const float dkx = (float)(1.0 / ntx / (double)dx);
    const int az = ntz / 2 + (ntz % 2), ax = ntx / 2 + (ntx % 2);
    const float kz = (float)(2 * kPi * (double)dkz * (double)(iz < az ? iz : ntz - iz));
    const float kx = (float)(2 * kPi * (double)dkx * (double)(ix < ax ? ix : ntx - ix));
    k[i] = (float)((double)kx * (double)kx + (double)kz * (double)kz);
}

// re[unit][iz][ix] (row pitch rp) = p1 of each unit's current step.
__global__ void psm_pack_kernel(const UnitDesc* __restrict__ units, int s, double* __restrict__ re, int rp,
                                int64_t ustride, int ntp)
{
    const int un = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntp) return;
    const UnitDesc& u = units[un];
    const int ls = s - u.start + u.s_off;  // the unit's local step
    if (ls < u.s_off || ls >= u.s_off + u.total) return;
    const StepInfo si = decode_step(u, ls);
    if (si.mode == kIdle) return;
    const FieldPtrs f = step_fields(u, si);
    const int iz = i / u.ntx, ix = i - iz * u.ntx;
    re[un * ustride + (int64_t)iz * rp + ix] = (double)f.p1[(int64_t)iz * u.pitch + ix];
}

// spectrum *= -k (both components; the float k is widened to double).  The
// spectrum holds ntz rows of nc = ntx/2 + 1 columns per unit; column ix of
// row iz takes k[iz][ix].
__global__ void psm_scale_kernel(double2* __restrict__ spec, const float* __restrict__ k, int ntx, int nc,
                                 int64_t per_unit, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = i % per_unit;
        const int iz = (int)(j / nc), ix = (int)(j - (int64_t)iz * nc);
        const double mk = (double)(-k[(int64_t)iz * ntx + ix]);
        double2 c = spec[i];
        c.x = mk * c.x;
        c.y = mk * c.y;
        spec[i] = c;
    }
}

// Everything after the Laplacian, one thread per padded point of a unit.
__global__ void psm_point_kernel(const UnitDesc* __restrict__ units, int s, double* __restrict__ re_all,
                                 int rp, int64_t ustride, int ntp)
{
    const int un = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntp) return;
    const UnitDesc& u = units[un];
    const int ls = s - u.start + u.s_off;  // the unit's local step
    if (ls < u.s_off || ls >= u.s_off + u.total) return;
    const StepInfo si = decode_step(u, ls);
    if (si.mode == kIdle) return;
    const FieldPtrs f = step_fields(u, si);
    const int64_t margin = (int64_t)kPadRows * u.pitch;
    const StepCtx c = make_step(u, si.mode, si.it, si.seg, f.out, margin);
    const PointConsts kc = point_consts(u);
    const int ntx = kc.ntx, ntz = kc.ntz, pml = kc.pml, pitch4 = kc.pitch4;
    const int iz = i / ntx, ix = i - iz * ntx;
    // byte offset of (iz, ix) from the first margin row, split so that the
    // inward reads (up to 4 rows or columns either way) stay >= 0
    const int ro = 4 * pitch4 + 16;
    const unsigned vo = (unsigned)((iz + kPadRows) * pitch4 + ix * 4 - ro);
    const Rsrc rp1 = make_rsrc(f.p1 - margin);
    const Rsrc rp0 = make_rsrc(f.p0 - margin);
    const Rsrc rv = make_rsrc(u.vel - margin);
    const float p1c = bload(rp1, vo, ro), p0c = bload(rp0, vo, ro), v = bload(rv, vo, ro);
    double* const re_pt = re_all + un * ustride + (int64_t)iz * rp + ix;
    const double re = *re_pt;
    const float vdt = v * kc.dt;
    float val = (float)((double)(2.f * p1c - p0c) + (double)(vdt * vdt) * re / (double)ntp);

    const bool top = iz < pml, bot = iz >= ntz - pml, lef = ix < pml, rig = ix >= ntx - pml;
    if (top || bot || lef || rig) {
        const float sc = div1_const(v * kc.dt, kc.dx, kc.rdx, kc.hi, kc.ydx);
        const float t11 = ((2.f - sc) * (1.f - sc)) / 2.f;
        const float t12 = sc * (2.f - sc);
        const float t13 = (sc * (sc - 1.f)) / 2.f;
        const Rsrc rw = make_rsrc(u.wtab);
        if (top)
            val = abc_blend(bload(rw, (unsigned)iz * 4u, 0),
                            abc_extrapolate(p1c, bload(rp1, vo, ro + pitch4), bload(rp1, vo, ro + 2 * pitch4), rp0,
                                            vo, ro, pitch4, t11, t12, t13), val);
        if (bot)
            val = abc_blend(bload(rw, (unsigned)(ntz - iz - 1) * 4u, 0),
                            abc_extrapolate(p1c, bload(rp1, vo, ro - pitch4), bload(rp1, vo, ro - 2 * pitch4), rp0,
                                            vo, ro, -pitch4, t11, t12, t13), val);
        if (lef)
            val = abc_blend(bload(rw, (unsigned)ix * 4u, 0),
                            abc_extrapolate(p1c, bload(rp1, vo, ro + 4), bload(rp1, vo, ro + 8), rp0, vo, ro, 4, t11,
                                            t12, t13), val);
        if (rig)
            val = abc_blend(bload(rw, (unsigned)(ntx - ix - 1) * 4u, 0),
                            abc_extrapolate(p1c, bload(rp1, vo, ro - 4), bload(rp1, vo, ro - 8), rp0, vo, ro, -4, t11,
                                            t12, t13), val);
    }
    const float dt = kc.dt;
    const bool rec_row = iz == u.r_iz;
    if (c.mode != kAdj) {
        if (iz == u.s_iz && ix == u.s_ix) val = val + G(u.wav)[c.it] * dt * dt * v * v;
    } else if (rec_row) {
        const int d = ix - u.r_ix0;
        if (d >= 0 && d % u.dr == 0 && d / u.dr < u.nr) {
            const int ri = c.it * u.nr + d / u.dr;
            const float res = kc.run_fwi ? G(u.syn)[ri] - G(u.obs)[ri] : G(u.obs)[ri];
            val = val + res * dt * dt * v * v;
        }
    }
    const bool inner = iz >= pml && iz < ntz - pml && ix >= pml && ix < ntx - pml;
    const unsigned gvo = inner ? (unsigned)(((iz - pml) * u.nx + (ix - pml)) * 4) : 0u;
    store_point(u, kc, c, val, p0c, p1c, v, ix, vo, ro, gvo, 0, make_rsrc(u.grad), rec_row, inner);
    *re_pt = (double)val;  // the next step's transform input (this thread's own element)
}

hipError_t launch_psm_wavenumber(hipStream_t st, float* k, int ntx, int ntz, float dx, float dz)
{
    const int n = ntx * ntz;
    hipLaunchKernelGGL(psm_wavenumber_kernel, dim3((n + 255) / 256), dim3(256), 0, st, k, ntx, ntz, dx, dz);
    return hipGetLastError();
}

// In-place real/complex buffer of one unit: ntz rows of rp = 2 (ntx/2 + 1)
// doubles (real view) = ntx/2 + 1 complex values (spectrum view).
static inline int psm_rp(int ntx) { return 2 * (ntx / 2 + 1); }

hipError_t launch_psm_pack(hipStream_t st, const UnitDesc* units, int nunit, int s, double* buf, int ntx, int ntz)
{
    const int ntp = ntx * ntz, rp = psm_rp(ntx);
    hipLaunchKernelGGL(psm_pack_kernel, dim3((ntp + 255) / 256, nunit), dim3(256), 0, st, units, s, buf, rp,
                       (int64_t)ntz * rp, ntp);
    return hipGetLastError();
}

hipError_t launch_psm_scale(hipStream_t st, double* buf, const float* k, int ntx, int ntz, int nunit)
{
    const int nc = ntx / 2 + 1;
    const int64_t per_unit = (int64_t)ntz * nc, n = per_unit * nunit;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(psm_scale_kernel, dim3(blocks), dim3(256), 0, st, reinterpret_cast<double2*>(buf), k, ntx,
                       nc, per_unit, n);
    return hipGetLastError();
}

hipError_t launch_psm_point(hipStream_t st, const UnitDesc* units, int nunit, int s, double* buf, int ntx,
                            int ntz)
{
    const int ntp = ntx * ntz, rp = psm_rp(ntx);
    hipLaunchKernelGGL(psm_point_kernel, dim3((ntp + 255) / 256, nunit), dim3(256), 0, st, units, s, buf, rp,
                       (int64_t)ntz * rp, ntp);
    return hipGetLastError();
}

}  // namespace gpvi
