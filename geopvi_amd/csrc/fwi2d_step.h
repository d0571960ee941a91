// The fused time step of the 2-D acoustic FDTD solver for gfx950 (see
// fwi2d_device.h for the data layout).
//
// Work decomposition: every wavefront independently owns one "strip
// segment" of one unit -- 64 consecutive columns (one per lane) times `zseg`
// consecutive rows -- and marches down it in blocks of kRowBlock rows.  Each
// lane keeps p1 of its own column around the current block in registers
// (the z-neighbours of the stencil), the current row with its x-halo is
// staged in a wave-private LDS row (the x-neighbours), and every load of the
// next block (p1, x-halo, p0, velocity) is issued before the current block
// is computed.  Field buffers carry margin rows and a 64-column pitch, so
// these loads need no bounds checks.  No workgroup barrier is needed, so
// waves never wait for each other.
//
// Arithmetic follows the reference's expression shapes and float/double
// typing (geopvi/forward/fwi2d/fdtd2D_modelling.c); the library is compiled
// with -ffp-contract=off, so a point's update rounds exactly as the
// reference's float build does.
#pragma once

#include "fwi2d_device.h"

namespace gpvi {

// Global-address-space views: loads/stores through these compile to
// global_load/global_store with a 32-bit vector offset.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* G(const T* p)
{
    return (const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* G(T* p)
{
    return (__attribute__((address_space(1))) T*)p;
}

// ---------------------------------------------------------------------------
// Division by a per-job constant b with y = RN(1/b): q0 = a*y, r = a - b*q0
// (exact through the FMA), q = RN(q0 + r*y) is the correctly rounded a/b
// (Markstein's theorem) whenever no intermediate underflows or overflows.
// That holds for 2^-16 <= |b| <= 2^16 and a == 0 or 2^-80 <= |a| < 2^80;
// other lanes take the IEEE division.  The host checks the divisor range and
// verifies the identity over a full binade before it sets UnitDesc::fastdiv.
__device__ __forceinline__ bool cdiv_in_range(float a)
{
    const float aa = __builtin_fabsf(a);
    return (aa < 0x1p80f) && (aa >= 0x1p-80f || aa == 0.f);
}

__device__ __forceinline__ float cdiv_fast(float a, float b, float y)
{
    const float q0 = a * y;
    const float r = __builtin_fmaf(-q0, b, a);
    return __builtin_fmaf(r, y, q0);
}

// (a / b) / b, as the reference writes d2p/dx/dx and (.)/dt/dt.
__device__ __forceinline__ float div2_const(float a, float b, float y, bool fast)
{
    if (fast) {
        const float q = cdiv_fast(cdiv_fast(a, b, y), b, y);
        if (__builtin_expect(cdiv_in_range(a), 1)) return q;
    }
    return (a / b) / b;
}

__device__ __forceinline__ float div1_const(float a, float b, float y, bool fast)
{
    if (fast) {
        const float q = cdiv_fast(a, b, y);
        if (__builtin_expect(cdiv_in_range(a), 1)) return q;
    }
    return a / b;
}

// Order-2 transmitting extrapolation (fdtd2D_modelling.c:840-846): q1..q3
// are p1 at 0, 1, 2 cells inward, a0..a4 are p0 at 0..4 cells inward.
__device__ __forceinline__ float abc_extrapolate(float q1, float q2, float q3, const float* p0, int ip, int d0,
                                                 float t11, float t12, float t13)
{
    const auto g = G(p0);
    const float a0 = g[ip], a1 = g[ip + d0], a2 = g[ip + 2 * d0], a3 = g[ip + 3 * d0], a4 = g[ip + 4 * d0];
    float h = 2.f * t11 * q1;
    h = h + 2.f * t12 * q2;
    h = h + 2.f * t13 * q3;
    h = h - t11 * t11 * a0;
    h = h - 2.f * t11 * t12 * a1;
    h = h - (t12 * t12 + 2.f * t11 * t13) * a2;
    h = h - 2.f * t12 * t13 * a3;
    h = h - t13 * t13 * a4;
    return h;
}

// w*H + (1.0 - w)*U: float product, double blend, float store (fdtd.c:840-847).
__device__ __forceinline__ float abc_blend(float w, float h, float u)
{
    return (float)((double)(w * h) + (1.0 - (double)w) * (double)u);
}

// Per-step, per-unit quantities shared by every point a wave updates.
struct StepCtx {
    const float* p0;  // field two steps back (UnitDesc::zero before the pass starts)
    const float* p1;  // field one step back
    float* out;
    float* save;      // this step's store/imaging slice (nx*nz), or nullptr
    float* ck0;       // checkpoint slices written this step, or nullptr
    float* ck1;
    int it, mode;
    bool first_image;
};

__device__ __forceinline__ StepCtx make_ctx(const UnitDesc& u, const StepInfo& si)
{
    const int64_t fe = u.field_elems;
    const float* pre0 = u.zero;
    const float* pre1 = u.zero;
    float* base = u.fwd;
    if (si.mode == kAdj) {
        base = u.adj;
    } else if (si.mode == kRec && si.seg > 0) {
        pre0 = u.ck + (int64_t)(2 * si.seg) * fe;
        pre1 = pre0 + fe;
    }
    const FieldPtrs f = rotate_ptrs(base, fe, pre0, pre1, si.q);
    StepCtx c;
    c.p0 = f.p0;
    c.p1 = f.p1;
    c.out = f.out;
    c.it = si.it;
    c.mode = si.mode;
    const bool fwd_like = si.mode != kAdj;
    const bool on_k = (si.it % u.k) == 0;
    const bool store = fwd_like && u.store && on_k && (si.mode == kRec || u.nseg == 1);
    const bool image = !fwd_like && on_k;
    const int slot = (u.nseg == 1) ? si.it / u.k : (si.it - si.seg * u.ckpt) / u.k;
    c.save = (store || image) ? u.save + (int64_t)slot * u.nx * u.nz : nullptr;
    const bool ckpt = (si.mode == kFwd) && (u.nseg > 1) && (si.it > 0) && (si.it % u.ckpt == 0);
    if (ckpt) {
        const int j = si.it / u.ckpt;
        c.ck0 = u.ck + (int64_t)(2 * j) * fe;
        c.ck1 = c.ck0 + fe;
    } else {
        c.ck0 = c.ck1 = nullptr;
    }
    c.first_image = u.grad_init && (si.it == ((u.nt - 1) / u.k) * u.k);
    return c;
}

// Everything after the stencil and ABC for one point: source or residual
// injection, checkpoint, wavefield store / record (forward-like steps) or
// imaging condition (adjoint steps), and the output store.
__device__ __forceinline__ void finish_point(const UnitDesc& u, const StepCtx& c, float val, float p0c, float p1c,
                                             float v, int ix, int iz, int ip, bool rec_row, bool src_row,
                                             bool inner)
{
    const float dt = u.dt;
    int ir = -1;
    if (rec_row) {
        const int d = ix - u.r_ix0;
        if (d >= 0 && d % u.dr == 0 && d / u.dr < u.nr) ir = d / u.dr;
    }
    const int gi = (iz - u.pml) * u.nx + (ix - u.pml);
    if (c.mode != kAdj) {
        // S3: source after the ABC, then checkpoint, store and record.
        if (src_row && ix == u.s_ix) val = val + G(u.wav)[c.it] * dt * dt * v * v;
        if (c.ck0) {
            G(c.ck0)[ip] = p0c;
            G(c.ck1)[ip] = p1c;
        }
        if (c.save && inner) {
            float sv = val;
            if (u.run_fwi) sv = div2_const((val - 2.f * p1c) + p0c, dt, u.rdt, u.fastdiv != 0);
            G(c.save)[gi] = sv;
        }
        if (c.mode == kFwd && ir >= 0) G(u.syn)[c.it * u.nr + ir] = val;
    } else {
        // S5: residual injection, then the imaging condition.
        if (ir >= 0) {
            const int ri = c.it * u.nr + ir;
            const float res = u.run_fwi ? G(u.syn)[ri] - G(u.obs)[ri] : G(u.obs)[ri];
            val = val + res * dt * dt * v * v;
        }
        if (c.save && inner) {
            const float sv = G(c.save)[gi];
            const float gprev = c.first_image ? 0.f : G(u.grad)[gi];
            float gnew;
            if (u.run_fwi) {
                const double vd = (double)v;
                const float v3 = (float)(vd * vd * vd);
                const double term = -2.0 / (double)v3 * (double)sv * (double)val * (double)u.k * (double)dt;
                gnew = (float)((double)gprev + term);
            } else {
                gnew = gprev + sv * val * (float)u.k * dt;
            }
            G(u.grad)[gi] = gnew;
        }
    }
    G(c.out)[ip] = val;
}


// One wave: one strip segment of one unit at local step s.  `row` is the
// wave's private LDS row of 64 + 2*(LC-1) floats.
template <int LC>
__device__ __forceinline__ void process_strip(const UnitDesc& u, int s, int item, float* row)
{
    constexpr int H = LC - 1;
    constexpr int R = kRowBlock;
    constexpr int NWIN = 2 * H + R;  // p1 rows a block needs in each lane's column
    const StepInfo si = decode_step(u, s);
    if (si.mode == kIdle) return;
    const StepCtx c = make_ctx(u, si);

    const int ntx = u.ntx, ntz = u.ntz, pitch = u.pitch, pml = u.pml;
    const int seg = item / u.tiles_x;
    const int x0 = (item - seg * u.tiles_x) * 64;
    const int zb = seg * u.zseg;
    const int ze = min(zb + u.zseg, ntz);
    const int lane = threadIdx.x & 63;
    const int ix = x0 + lane;
    const bool lane_ok = ix < ntx;

    // Halo lanes: lanes 0..H-1 also carry columns x0-H.., lanes H..2H-1
    // columns x0+64..; their LDS slots are lane and 64+lane.  Other lanes
    // load their own column again (the value is not used).
    const bool halo_lane = lane < 2 * H;
    const int hcol = lane < H ? x0 - H + lane : (halo_lane ? x0 + 64 + (lane - H) : ix);
    const int hslot = lane < H ? lane : 64 + lane;

    const bool strip_inside = x0 >= pml && x0 + 64 <= ntx - pml;
    const bool lane_interior_x = ix >= LC && ix < ntx - LC;
    const bool lane_inner_x = ix >= pml && ix < ntx - pml;
    const bool lane_band_x = (ix >= LC && ix < pml) || (ix >= ntx - pml && ix < ntx - LC);
    const bool fast = u.fastdiv != 0;
    const auto P1 = G(c.p1);
    const auto P0 = G(c.p0);
    const auto V = G(u.vel);

    float win[NWIN];  // p1 at rows z0-H .. z0+R+H-1 of this lane's column
    float hal[R], p0r[R], vr[R];
#pragma unroll
    for (int j = 0; j < NWIN; ++j) win[j] = P1[(zb - H + j) * pitch + ix];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        hal[r] = P1[(zb + r) * pitch + hcol];
        p0r[r] = P0[(zb + r) * pitch + ix];
        vr[r] = V[(zb + r) * pitch + ix];
    }

    for (int z0 = zb; z0 < ze; z0 += R) {
        // Issue every load of the next block before computing this one.
        const int z1 = z0 + R;
        float nwin[R], nhal[R], np0[R], nv[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            nwin[r] = P1[(z1 + H + r) * pitch + ix];
            nhal[r] = P1[(z1 + r) * pitch + hcol];
            np0[r] = P0[(z1 + r) * pitch + ix];
            nv[r] = V[(z1 + r) * pitch + ix];
        }

#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int iz = z0 + r;
            if (iz >= ze) break;
            const float p1c = win[H + r];
            const float p0c = p0r[r];
            const float v = vr[r];
            // Stage the row (centre + x-halo) in the wave's LDS row.
            row[lane + H] = p1c;
            if (halo_lane) row[hslot] = hal[r];
            __builtin_amdgcn_wave_barrier();
            const float* sp = row + lane + H;

            const bool row_interior = iz >= LC && iz < ntz - LC;
            const bool row_inner = iz >= pml && iz < ntz - pml;
            const bool rec_row = iz == u.r_iz;
            const bool src_row = iz == u.s_iz && c.mode != kAdj;
            const int ip = iz * pitch + ix;

            float val = p1c;
            if (strip_inside && row_inner) {
                // Inner grid, away from every absorbing band.
                float dxx = u.fd[0] * p1c;
#pragma unroll
                for (int m = 1; m < LC; ++m) dxx = dxx + u.fd[m] * (sp[m] + sp[-m]);
                float dzz = u.fd[0] * p1c;
#pragma unroll
                for (int m = 1; m < LC; ++m) dzz = dzz + u.fd[m] * (win[H + r + m] + win[H + r - m]);
                const float vdt = v * u.dt;
                const float lap = div2_const(dxx, u.dx, u.rdx, fast) + div2_const(dzz, u.dz, u.rdz, fast);
                val = (2.f * p1c - p0c) + (vdt * vdt) * lap;
                finish_point(u, c, val, p0c, p1c, v, ix, iz, ip, rec_row, src_row, true);
            } else if (lane_ok) {
                if (row_interior && lane_interior_x) {
                    float dxx = u.fd[0] * p1c;
#pragma unroll
                    for (int m = 1; m < LC; ++m) dxx = dxx + u.fd[m] * (sp[m] + sp[-m]);
                    float dzz = u.fd[0] * p1c;
#pragma unroll
                    for (int m = 1; m < LC; ++m) dzz = dzz + u.fd[m] * (win[H + r + m] + win[H + r - m]);
                    const float vdt = v * u.dt;
                    const float lap = div2_const(dxx, u.dx, u.rdx, fast) + div2_const(dzz, u.dz, u.rdz, fast);
                    val = (2.f * p1c - p0c) + (vdt * vdt) * lap;
                }
                // S2: top, bottom, left, right, in that order.
                const bool top = iz >= LC && iz < pml;
                const bool bot = iz >= ntz - pml && iz < ntz - LC;
                if (top || bot || lane_band_x) {
                    const float sc = div1_const(v * u.dt, u.dx, u.rdx, fast);
                    const float t11 = ((2.f - sc) * (1.f - sc)) / 2.f;
                    const float t12 = sc * (2.f - sc);
                    const float t13 = (sc * (sc - 1.f)) / 2.f;
                    const auto wt = G(u.wtab);
                    if (top)
                        val = abc_blend(wt[iz - LC],
                                        abc_extrapolate(p1c, win[H + r + 1], win[H + r + 2], c.p0, ip, pitch, t11,
                                                        t12, t13),
                                        val);
                    if (bot)
                        val = abc_blend(wt[ntz - LC - iz - 1],
                                        abc_extrapolate(p1c, win[H + r - 1], win[H + r - 2], c.p0, ip, -pitch, t11,
                                                        t12, t13),
                                        val);
                    if (ix >= LC && ix < pml)
                        val = abc_blend(wt[ix - LC], abc_extrapolate(p1c, sp[1], sp[2], c.p0, ip, 1, t11, t12, t13),
                                        val);
                    if (ix >= ntx - pml && ix < ntx - LC)
                        val = abc_blend(wt[ntx - LC - ix - 1],
                                        abc_extrapolate(p1c, sp[-1], sp[-2], c.p0, ip, -1, t11, t12, t13), val);
                }
                finish_point(u, c, val, p0c, p1c, v, ix, iz, ip, rec_row, src_row, row_inner && lane_inner_x);
            }
            __builtin_amdgcn_wave_barrier();
        }

        // Slide the window to the next block.
#pragma unroll
        for (int j = 0; j < 2 * H; ++j) win[j] = win[j + R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            win[2 * H + r] = nwin[r];
            hal[r] = nhal[r];
            p0r[r] = np0[r];
            vr[r] = nv[r];
        }
    }
}

}  // namespace gpvi
