// Pseudo-spectral Laplacian path (laplace_slover = 1) of the FWI-2D solver
// for gfx950: geopvi/forward/fwi2d/fdtd2D_modelling.c with Lc = 0
// (pml = pml0, fdtd.c:75-85) and pstd_2d_calculate_p (fdtd.c:783-830) in
// place of the stencil.  One time step of every unit of a batch, when both
// padded lengths factor into 2, 3, 5, 7, 11 and 13 and fit the LDS tiling
// (fwi2d_spectral.h):
//
//   rows forward    row pairs of p1 -> half spectra    (fdtd.c:791-803)
//   columns         column transforms, times -|k|^2, back
//                                                     (fdtd.c:805-817)
//   rows backward   half spectra -> rows, then per point
//                   p2 = 2 p1 - p0 + (v dt)^2 Re / (ntx ntz), the four
//                   ordered absorbing sweeps over the pml0-wide bands, source
//                   or residual injection, store/records or imaging
//                                                     (fdtd.c:819-826, 833-978);
//                   within a pass (forward, a segment's recompute, an adjoint
//                   stretch) also the next step's rows forward on the new
//                   p2 rows, so rows forward runs alone only at a pass's
//                   first step
//
// and otherwise (or with GPVI_PSM_HIPFFT set) through hipFFT:
//
//   pack    p1 -> double field; only at the first step of a pass: within a
//           pass the point kernel of the previous step leaves its result there
//   FFT     batched 2-D forward transform         (hipFFT, D2Z)
//   scale   multiply by -|k|^2
//   IFFT    batched 2-D inverse (unnormalised)    (hipFFT, Z2D)
//   point   as the rows-backward kernel's point update
//
// The reference transforms the real field as a complex one (fftw c2c) and
// keeps the real part of the inverse.  The spectrum of a real field is
// Hermitian and |k|^2 is even in both axes (kx(ix) = kx(ntx - ix)), so the
// scaled spectrum stays Hermitian and its inverse is real: real-to-complex /
// complex-to-real transforms compute the same values with half the work;
// only the rounding differs, as between any two FFT libraries.  Transforms
// are double precision (FFTW's fftw_complex is double); a unit's buffer
// holds ntz rows of nc = ntx/2 + 1 complex values (for hipFFT, in place:
// real rows padded to 2 nc doubles).  Every other operation keeps the
// reference's float/double typing.
#include "fwi2d_kernels.h"
#include "fwi2d_step.h"
#include "fwi2d_spectral.h"

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

// p1, p0 and v at point (iz, ix) of a unit.
struct PsmCore {
    float p1, p0, v;
};
__device__ inline unsigned psm_vo(const PointConsts& kc, int iz, int ix, int ro)
{
    return (unsigned)((iz + kPadRows) * kc.pitch4 + ix * 4 - ro);
}
__device__ inline PsmCore psm_core(const UnitDesc& u, const FieldPtrs& f, const PointConsts& kc, int iz, int ix)
{
    const int64_t margin = (int64_t)kPadRows * u.pitch;
    const int ro = 4 * kc.pitch4 + 16;
    const unsigned vo = psm_vo(kc, iz, ix, ro);
    PsmCore r;
    r.p1 = bload(make_rsrc(f.p1 - margin), vo, ro);
    r.p0 = bload(make_rsrc(f.p0 - margin), vo, ro);
    r.v = bload(make_rsrc(u.vel - margin), vo, ro);
    return r;
}

// Everything after the Laplacian at point (iz, ix) of a unit, from its p1,
// p0 and v: the update p2 = 2 p1 - p0 + (v dt)^2 re / div (re / div = the
// Laplacian), the four ordered absorbing blends over the pml0-wide bands,
// injection, store or imaging and records.  Returns p2.
__device__ inline float psm_point(const UnitDesc& u, const FieldPtrs& f, const StepCtx& c, const PointConsts& kc,
                                  int iz, int ix, double re, double div, PsmCore core)
{
    const int ntx = kc.ntx, ntz = kc.ntz, pml = kc.pml, pitch4 = kc.pitch4;
    const int64_t margin = (int64_t)kPadRows * u.pitch;
    // byte offset of (iz, ix) from the first margin row, split so that the
    // inward reads (up to 4 rows or columns either way) stay >= 0
    const int ro = 4 * pitch4 + 16;
    const unsigned vo = psm_vo(kc, iz, ix, ro);
    const Rsrc rp1 = make_rsrc(f.p1 - margin);
    const Rsrc rp0 = make_rsrc(f.p0 - margin);
    const float p1c = core.p1, p0c = core.p0, v = core.v;
    const float vdt = v * kc.dt;
    float val = (float)((double)(2.f * p1c - p0c) + (double)(vdt * vdt) * re / div);

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
    return val;
}

// The unit of a block at batch step s and its step (false: idle at s).
__device__ inline bool psm_unit_step(const UnitDesc& u, int s, StepInfo* si)
{
    const int ls = s - u.start + u.s_off;  // the unit's local step
    if (ls < u.s_off || ls >= u.s_off + u.total) return false;
    *si = decode_step(u, ls);
    return si->mode != kIdle;
}

// After the library transforms: one thread per padded point of a unit.
__global__ void psm_point_kernel(const UnitDesc* __restrict__ units, int s, double* __restrict__ re_all,
                                 int rp, int64_t ustride, int ntp)
{
    const int un = blockIdx.y;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntp) return;
    const UnitDesc& u = units[un];
    StepInfo si;
    if (!psm_unit_step(u, s, &si)) return;
    const FieldPtrs f = step_fields(u, si);
    const int64_t margin = (int64_t)kPadRows * u.pitch;
    const StepCtx c = make_step(u, si.mode, si.it, si.seg, f.out, margin);
    const PointConsts kc = point_consts(u);
    const int iz = i / kc.ntx, ix = i - iz * kc.ntx;
    double* const re_pt = re_all + un * ustride + (int64_t)iz * rp + ix;
    const float val = psm_point(u, f, c, kc, iz, ix, *re_pt, (double)ntp, psm_core(u, f, kc, iz, ix));
    *re_pt = (double)val;  // the next step's transform input (this thread's own element)
}

// ---------------------------------------------------------------------------
// Hand-written transforms (fwi2d_spectral.h) over a per-unit half spectrum
// of ntz x nc complex doubles (nc = ntx/2 + 1): rows forward (first step of
// a pass only), columns, rows backward + point update (+ the next step's
// rows forward within a pass).
using spec::Cx;

extern __shared__ __align__(16) unsigned char psm_lds[];

// The twiddle tables of a job, (cos, sin)(2 pi m / n) for m < n, n = ntx
// then n = ntz; made once per call, copied to LDS by every workgroup.
__global__ void psm_twiddle_kernel(Cx* __restrict__ tw, int ntx, int ntz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntx + ntz) return;
    const int n = i < ntx ? ntx : ntz, m = i < ntx ? i : i - ntx;
    double sn, cs;
    sincospi(2.0 * (double)m / (double)n, &sn, &cs);
    tw[i] = spec::cx(cs, sn);
}

__device__ inline void psm_table(Cx* T, const Cx* __restrict__ tw, int n)
{
    for (int m = threadIdx.x; m < spec::table_len(n); m += blockDim.x) T[m] = tw[m];
}

// The plan's passes over nseq sequences, x -> y -> x ... (a barrier after
// each); returns the buffer holding the result.  Expects a barrier before.
template <int MaxR>
__device__ inline Cx* psm_transform(Cx* x, Cx* y, int nseq, const spec::Plan& p, const Cx* T, int sgn)
{
    int ns = 1;
    for (int i = 0; i < p.npass; ++i) {
        spec::run_pass<MaxR>(p.radix[i], x, y, nseq, p.n, ns, T, sgn, threadIdx.x, blockDim.x);
        __syncthreads();
        ns *= p.radix[i];
        Cx* t = x;
        x = y;
        y = t;
    }
    return x;
}

// Rows forward: block (g, unit) transforms row pairs [g np, (g + 1) np) of
// p1 and writes the half spectra of their rows.
template <int MaxR>
__global__ void __launch_bounds__(spec::kThreads)
psm_rows_fwd_kernel(const UnitDesc* __restrict__ units, int s, Cx* __restrict__ spec_all, int64_t ustride,
                    const Cx* __restrict__ tw, spec::Plan px, int ntz, int np)
{
    const UnitDesc& u = units[blockIdx.y];
    StepInfo si;
    if (!psm_unit_step(u, s, &si)) return;  // the whole block
    const FieldPtrs f = step_fields(u, si);
    const int ntx = px.n, nc = ntx / 2 + 1, pair0 = blockIdx.x * np;
    Cx* const sp = spec_all + blockIdx.y * ustride;
    Cx* T = reinterpret_cast<Cx*>(psm_lds);
    Cx* X = T + spec::table_len(ntx);
    Cx* Y = X + np * ntx;
    psm_table(T, tw, ntx);
    const float* p1 = f.p1;
    const int pitch = u.pitch;
    spec::rows_load(X, np, ntx, ntz, pair0, [&](int iz, int ix) { return (double)p1[(int64_t)iz * pitch + ix]; },
                    threadIdx.x, blockDim.x);
    __syncthreads();
    const Cx* Z = psm_transform<MaxR>(X, Y, np, px, T, +1);
    spec::rows_unpack(Z, np, ntx, ntz, nc, pair0, [&](int iz, int fq, Cx v) { sp[(int64_t)iz * nc + fq] = v; },
                      threadIdx.x, blockDim.x);
}

// Columns: block (g, unit) transforms spectrum columns [g ncol, (g + 1) ncol),
// scales them by -|k|^2 and transforms them back.
template <int MaxR>
__global__ void __launch_bounds__(spec::kThreads)
psm_cols_kernel(const UnitDesc* __restrict__ units, int s, Cx* __restrict__ spec_all, int64_t ustride,
                const float* __restrict__ k, const Cx* __restrict__ tw, spec::Plan pz, int ntx, int ncol)
{
    // Consecutive workgroups go to different XCDs (each with its own L2),
    // and neighbouring column tiles share cache lines: give every XCD a
    // contiguous run of tiles (a permutation of the first 8 * floor(nb / 8)
    // workgroups; the rest keep their place).
    const int nb = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x;
    const int full = nb / 8 * 8;
    const int lb = b < full ? (b % 8) * (nb / 8) + b / 8 : b;
    const int unit = lb / gridDim.x, tile = lb - unit * gridDim.x;
    const UnitDesc& u = units[unit];
    StepInfo si;
    if (!psm_unit_step(u, s, &si)) return;  // the whole block
    const int ntz = pz.n, nc = ntx / 2 + 1, f0 = tile * ncol;
    Cx* const sp = spec_all + unit * ustride;
    Cx* T = reinterpret_cast<Cx*>(psm_lds);
    Cx* X = T + spec::table_len(ntz);
    Cx* Y = X + ncol * ntz;
    psm_table(T, tw + ntx, ntz);
    spec::cols_load(X, ncol, ntz, nc, f0, [&](int iz, int fq) { return sp[(int64_t)iz * nc + fq]; }, threadIdx.x,
                    blockDim.x);
    __syncthreads();
    Cx* Z = psm_transform<MaxR>(X, Y, ncol, pz, T, +1);
    spec::cols_scale(Z, ncol, ntz, nc, f0, [&](int iz, int fq) { return k[(int64_t)iz * ntx + fq]; }, threadIdx.x,
                     blockDim.x);
    __syncthreads();
    const Cx* W = psm_transform<MaxR>(Z, Z == X ? Y : X, ncol, pz, T, -1);
    spec::cols_store(W, ncol, ntz, nc, f0, [&](int iz, int fq, Cx v) { sp[(int64_t)iz * nc + fq] = v; },
                     threadIdx.x, blockDim.x);
}

// Rows backward and the point update: block (g, unit) transforms row pairs
// [g np, (g + 1) np) of the scaled spectrum back and finishes their points.
// chain (the next step continues this pass, so its p1 is this step's p2):
// the block then also does the next step's rows-forward work on the p2 rows
// it has just made, from LDS, and writes their half spectra over the ones it
// read (each block reads and writes only its own rows).
template <int MaxR>
__global__ void __launch_bounds__(spec::kThreads)
psm_rows_inv_kernel(const UnitDesc* __restrict__ units, int s, Cx* __restrict__ spec_all, int64_t ustride,
                    const Cx* __restrict__ tw, spec::Plan px, int ntz, int np, int chain)
{
    const UnitDesc& u = units[blockIdx.y];
    StepInfo si;
    if (!psm_unit_step(u, s, &si)) return;  // the whole block
    const int ntx = px.n, nc = ntx / 2 + 1, pair0 = blockIdx.x * np;
    Cx* const sp = spec_all + blockIdx.y * ustride;
    Cx* T = reinterpret_cast<Cx*>(psm_lds);
    Cx* X = T + spec::table_len(ntx);
    Cx* Y = X + np * ntx;
    psm_table(T, tw, ntx);
    spec::rows_repack(X, np, ntx, ntz, pair0, [&](int iz, int fq) { return sp[(int64_t)iz * nc + fq]; },
                      threadIdx.x, blockDim.x);
    __syncthreads();
    const Cx* Z = psm_transform<MaxR>(X, Y, np, px, T, -1);
    Cx* const W = Z == X ? Y : X;   // the next step's rows (chain)
    const FieldPtrs f = step_fields(u, si);
    const int64_t margin = (int64_t)kPadRows * u.pitch;
    const StepCtx c = make_step(u, si.mode, si.it, si.seg, f.out, margin);
    const PointConsts kc = point_consts(u);
    const double ntp = (double)ntx * (double)ntz;
    // the points of the block's rows, spec::kChunk per thread at a time with
    // their p1, p0 and v loads in flight together
    const int total = 2 * np * ntx;
    for (int t0 = threadIdx.x; t0 < total; t0 += spec::kChunk * blockDim.x) {
        PsmCore core[spec::kChunk];
#pragma unroll
        for (int j = 0; j < spec::kChunk; ++j) {
            const int t = t0 + j * blockDim.x;
            const int q = t / (2 * ntx), r = t - q * 2 * ntx;
            const int w = r >= ntx ? 1 : 0, ix = r - w * ntx;
            const int iz = 2 * (pair0 + q) + w;
            if (t < total && iz < ntz) core[j] = psm_core(u, f, kc, iz, ix);
        }
#pragma unroll
        for (int j = 0; j < spec::kChunk; ++j) {
            const int t = t0 + j * blockDim.x;
            const int q = t / (2 * ntx), r = t - q * 2 * ntx;
            const int w = r >= ntx ? 1 : 0, ix = r - w * ntx;
            const int iz = 2 * (pair0 + q) + w;
            if (t >= total) continue;
            double* const nxt = &W[q * ntx + ix].re + w;
            if (iz >= ntz) {
                if (chain) *nxt = 0.0;   // the missing row of an odd last pair
                continue;
            }
            const Cx z = Z[q * ntx + ix];
            const float val = psm_point(u, f, c, kc, iz, ix, w ? z.im : z.re, ntp, core[j]);
            if (chain) *nxt = (double)val;
        }
    }
    if (!chain) return;   // uniform over the launch
    __syncthreads();
    const Cx* F = psm_transform<MaxR>(W, W == X ? Y : X, np, px, T, +1);
    spec::rows_unpack(F, np, ntx, ntz, nc, pair0, [&](int iz, int fq, Cx v) { sp[(int64_t)iz * nc + fq] = v; },
                      threadIdx.x, blockDim.x);
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

static bool psm_spectral_plans(int ntx, int ntz, spec::Plan* px, spec::Plan* pz, int* np, int* ncol)
{
    if (!spec::make_plan(ntx, px) || !spec::make_plan(ntz, pz)) return false;
    *np = spec::pairs_per_group(ntx);
    *ncol = spec::cols_per_group(ntz);
    return *np > 0 && *ncol > 0;
}

bool psm_spectral_supported(int ntx, int ntz)
{
    spec::Plan px, pz;
    int np, ncol;
    return psm_spectral_plans(ntx, ntz, &px, &pz, &np, &ncol);
}

hipError_t launch_psm_twiddles(hipStream_t st, double* tw, int ntx, int ntz)
{
    const int n = ntx + ntz;
    hipLaunchKernelGGL(psm_twiddle_kernel, dim3((n + 255) / 256), dim3(256), 0, st, reinterpret_cast<Cx*>(tw), ntx,
                       ntz);
    return hipGetLastError();
}

hipError_t launch_psm_spectral(hipStream_t st, const UnitDesc* units, int nunit, int s, double* buf,
                               const float* k, const double* twd, int ntx, int ntz, bool first, bool chain)
{
    spec::Plan px, pz;
    int np, ncol;
    if (!psm_spectral_plans(ntx, ntz, &px, &pz, &np, &ncol)) return hipErrorInvalidValue;
    const int nc = ntx / 2 + 1;
    // the buffer holds ntz * 2 nc doubles per unit = ntz * nc complex values
    const int64_t ustride = (int64_t)ntz * nc;
    Cx* sp = reinterpret_cast<Cx*>(buf);
    const Cx* tw = reinterpret_cast<const Cx*>(twd);
    const int row_groups = ((ntz + 1) / 2 + np - 1) / np, col_groups = (nc + ncol - 1) / ncol;
    const size_t lds_rows = sizeof(Cx) * ((size_t)spec::table_len(ntx) + (size_t)ntx * 2 * np);
    const size_t lds_cols = sizeof(Cx) * ((size_t)spec::table_len(ntz) + (size_t)ntz * 2 * ncol);
    // instantiations by largest radix: 7, 11 or 13
    const int mx = spec::max_radix(px), mz = spec::max_radix(pz);
    const dim3 rows(row_groups, nunit), cols(col_groups, nunit), blk(spec::kThreads);
    const int ch = chain ? 1 : 0;
    if (first) {
        if (mx <= 7)
            hipLaunchKernelGGL(psm_rows_fwd_kernel<7>, rows, blk, lds_rows, st, units, s, sp, ustride, tw, px, ntz, np);
        else if (mx == 11)
            hipLaunchKernelGGL(psm_rows_fwd_kernel<11>, rows, blk, lds_rows, st, units, s, sp, ustride, tw, px, ntz, np);
        else
            hipLaunchKernelGGL(psm_rows_fwd_kernel<13>, rows, blk, lds_rows, st, units, s, sp, ustride, tw, px, ntz, np);
    }
    if (mz <= 7)
        hipLaunchKernelGGL(psm_cols_kernel<7>, cols, blk, lds_cols, st, units, s, sp, ustride, k, tw, pz, ntx, ncol);
    else if (mz == 11)
        hipLaunchKernelGGL(psm_cols_kernel<11>, cols, blk, lds_cols, st, units, s, sp, ustride, k, tw, pz, ntx, ncol);
    else
        hipLaunchKernelGGL(psm_cols_kernel<13>, cols, blk, lds_cols, st, units, s, sp, ustride, k, tw, pz, ntx, ncol);
    if (mx <= 7)
        hipLaunchKernelGGL(psm_rows_inv_kernel<7>, rows, blk, lds_rows, st, units, s, sp, ustride, tw, px, ntz, np, ch);
    else if (mx == 11)
        hipLaunchKernelGGL(psm_rows_inv_kernel<11>, rows, blk, lds_rows, st, units, s, sp, ustride, tw, px, ntz, np,
                           ch);
    else
        hipLaunchKernelGGL(psm_rows_inv_kernel<13>, rows, blk, lds_rows, st, units, s, sp, ustride, tw, px, ntz, np,
                           ch);
    return hipGetLastError();
}

}  // namespace gpvi
