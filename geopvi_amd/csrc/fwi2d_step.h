// The fused time step of one tile of one unit (see fwi2d_device.h for the
// data layout).  Arithmetic follows the reference's expression shapes and
// float/double typing (geopvi/forward/fwi2d/fdtd2D_modelling.c), and the
// library is compiled with -ffp-contract=off, so that a point's update
// rounds exactly as the reference's float build does.
#pragma once

#include "fwi2d_device.h"

namespace gpvi {

// ABC taper weight w[i] = 1 - i/pml0, formed in double and stored as float
// (fdtd2D_modelling.c:227-230).
__device__ __forceinline__ float abc_weight(int i, int pml0)
{
    return (float)(1.0 - (double)i / (double)pml0);
}

// Order-2 transmitting extrapolation at one point, inward stride d1 in the
// LDS copy of p1 and d0 in global p0 (fdtd2D_modelling.c:840-846).
__device__ __forceinline__ float abc_extrapolate(const float* sp, int d1, const float* p0, int64_t ip,
                                                 int64_t d0, float t11, float t12, float t13)
{
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, a4 = 0.f;
    if (p0) {
        a0 = p0[ip]; a1 = p0[ip + d0]; a2 = p0[ip + 2 * d0]; a3 = p0[ip + 3 * d0]; a4 = p0[ip + 4 * d0];
    }
    float h = 2.f * t11 * sp[0];
    h = h + 2.f * t12 * sp[d1];
    h = h + 2.f * t13 * sp[2 * d1];
    h = h - t11 * t11 * a0;
    h = h - 2.f * t11 * t12 * a1;
    h = h - (t12 * t12 + 2.f * t11 * t13) * a2;
    h = h - 2.f * t12 * t13 * a3;
    h = h - t13 * t13 * a4;
    return h;
}

// w*H + (1.0 - w)*U: float product, double blend, float store.
__device__ __forceinline__ float abc_blend(float w, float h, float u)
{
    return (float)((double)(w * h) + (1.0 - (double)w) * (double)u);
}

template <int LC, int TX, int TZ, int NTH>
struct TileCfg {
    static constexpr int H = LC - 1;              // stencil halo
    static constexpr int W = TX + 2 * H;          // LDS row length
    static constexpr int HT = TZ + 2 * H;         // LDS rows
    static constexpr int kLds = W * HT;
    static constexpr int ZSTEP = NTH / TX;
    static_assert(NTH % TX == 0, "threads must cover whole tile rows");
    static_assert(TZ % ZSTEP == 0, "tile height must be a multiple of rows per pass");
};

template <int LC, int TX, int TZ, int NTH>
__device__ __forceinline__ void process_tile(const UnitDesc& u, int s, int tile, float* s_p1)
{
    using C = TileCfg<LC, TX, TZ, NTH>;
    constexpr int H = C::H, W = C::W;
    const StepInfo si = decode_step(u, s);
    if (si.mode == kIdle) return;

    const int tz = tile / u.tiles_x;
    const int tx = tile - tz * u.tiles_x;
    const int x0 = tx * TX, z0 = tz * TZ;
    const int ntx = u.ntx, ntz = u.ntz, pitch = u.pitch;
    const int64_t fe = u.field_elems;

    const float* pre0 = nullptr;
    const float* pre1 = nullptr;
    float* base = u.fwd;
    if (si.mode == kAdj) {
        base = u.adj;
    } else if (si.mode == kRec && si.seg > 0) {
        pre0 = u.ck + (int64_t)(2 * si.seg) * fe;
        pre1 = pre0 + fe;
    }
    const FieldPtrs f = rotate_ptrs(base, fe, pre0, pre1, si.q);

    // Stage p1 (tile + halo) in LDS; outside the padded grid reads as 0.
    if (f.p1) {
        for (int i = threadIdx.x; i < C::kLds; i += NTH) {
            const int r = i / W, c = i - r * W;
            const int gz = z0 - H + r, gx = x0 - H + c;
            float val = 0.f;
            if (gz >= 0 && gz < ntz && gx >= 0 && gx < ntx) val = f.p1[(int64_t)gz * pitch + gx];
            s_p1[i] = val;
        }
    } else {
        for (int i = threadIdx.x; i < C::kLds; i += NTH) s_p1[i] = 0.f;
    }
    __syncthreads();

    const int it = si.it;
    const float dt = u.dt, dx = u.dx, dz = u.dz;
    const int pml = u.pml, pml0 = u.pml0;
    const bool fwd_like = si.mode != kAdj;
    const bool do_store = fwd_like && u.store && (it % u.k == 0) && (si.mode == kRec || u.nseg == 1);
    const bool do_image = !fwd_like && (it % u.k == 0);
    const bool do_ckpt = (si.mode == kFwd) && (u.nseg > 1) && (it > 0) && (it % u.ckpt == 0);
    const int slot = (u.nseg == 1) ? it / u.k : (it - si.seg * u.ckpt) / u.k;
    const int ni = u.nx * u.nz;
    const bool first_image = u.grad_init && (it == ((u.nt - 1) / u.k) * u.k);

    const int lx = threadIdx.x % TX;
    const int lz0 = threadIdx.x / TX;
#pragma unroll
    for (int lz = lz0; lz < TZ; lz += C::ZSTEP) {
        const int ix = x0 + lx, iz = z0 + lz;
        if (ix >= ntx || iz >= ntz) continue;
        const int64_t ip = (int64_t)iz * pitch + ix;
        const float* sp = s_p1 + (lz + H) * W + (lx + H);
        const float p1c = sp[0];
        const float p0c = f.p0 ? f.p0[ip] : 0.f;
        const float v = u.vel[ip];

        // S1: interior update; points outside it start from p1.
        float val;
        if (iz >= LC && iz < ntz - LC && ix >= LC && ix < ntx - LC) {
            float dxx = u.fd[0] * p1c;
#pragma unroll
            for (int m = 1; m < LC; ++m) dxx = dxx + u.fd[m] * (sp[m] + sp[-m]);
            float dzz = u.fd[0] * p1c;
#pragma unroll
            for (int m = 1; m < LC; ++m) dzz = dzz + u.fd[m] * (sp[m * W] + sp[-m * W]);
            const float vdt = v * dt;
            val = (2.f * p1c - p0c) + (vdt * vdt) * ((dxx / dx) / dx + (dzz / dz) / dz);
        } else {
            val = p1c;
        }

        // S2: top, bottom, left, right, in that order.
        const bool top = iz >= LC && iz < pml;
        const bool bot = iz >= ntz - pml && iz < ntz - LC;
        const bool lef = ix >= LC && ix < pml;
        const bool rig = ix >= ntx - pml && ix < ntx - LC;
        if (top || bot || lef || rig) {
            const float sc = (v * dt) / dx;
            const float t11 = ((2.f - sc) * (1.f - sc)) / 2.f;
            const float t12 = sc * (2.f - sc);
            const float t13 = (sc * (sc - 1.f)) / 2.f;
            if (top)
                val = abc_blend(abc_weight(iz - LC, pml0),
                                abc_extrapolate(sp, W, f.p0, ip, pitch, t11, t12, t13), val);
            if (bot)
                val = abc_blend(abc_weight(ntz - LC - iz - 1, pml0),
                                abc_extrapolate(sp, -W, f.p0, ip, -(int64_t)pitch, t11, t12, t13), val);
            if (lef)
                val = abc_blend(abc_weight(ix - LC, pml0),
                                abc_extrapolate(sp, 1, f.p0, ip, 1, t11, t12, t13), val);
            if (rig)
                val = abc_blend(abc_weight(ntx - LC - ix - 1, pml0),
                                abc_extrapolate(sp, -1, f.p0, ip, -1, t11, t12, t13), val);
        }

        // Receiver index of this point, if any.
        int ir = -1;
        if (iz == u.r_iz) {
            const int d = ix - u.r_ix0;
            if (d >= 0 && d % u.dr == 0 && d / u.dr < u.nr) ir = d / u.dr;
        }

        if (fwd_like) {
            // S3: source after the ABC, then store and record.
            if (iz == u.s_iz && ix == u.s_ix) val = val + u.wav[it] * dt * dt * v * v;
            if (do_ckpt) {
                const int j = it / u.ckpt;
                u.ck[(int64_t)(2 * j) * fe + ip] = p0c;
                u.ck[(int64_t)(2 * j + 1) * fe + ip] = p1c;
            }
            if (do_store && iz >= pml && iz < ntz - pml && ix >= pml && ix < ntx - pml) {
                const int64_t gi = (int64_t)slot * ni + (int64_t)(iz - pml) * u.nx + (ix - pml);
                u.save[gi] = u.run_fwi ? ((val - 2.f * p1c) + p0c) / dt / dt : val;
            }
            if (si.mode == kFwd && ir >= 0) u.syn[(int64_t)it * u.nr + ir] = val;
        } else {
            // S5: residual injection, then the imaging condition.
            if (ir >= 0) {
                const int64_t ri = (int64_t)it * u.nr + ir;
                const float res = u.run_fwi ? u.syn[ri] - u.obs[ri] : u.obs[ri];
                val = val + res * dt * dt * v * v;
            }
            if (do_image && iz >= pml && iz < ntz - pml && ix >= pml && ix < ntx - pml) {
                const int gi = (iz - pml) * u.nx + (ix - pml);
                const float sv = u.save[(int64_t)slot * ni + gi];
                const float gprev = first_image ? 0.f : u.grad[gi];
                if (u.run_fwi) {
                    const double vd = (double)v;
                    const float v3 = (float)(vd * vd * vd);
                    const double term = -2.0 / (double)v3 * (double)sv * (double)val * (double)u.k * (double)dt;
                    u.grad[gi] = (float)((double)gprev + term);
                } else {
                    u.grad[gi] = gprev + sv * val * (float)u.k * dt;
                }
            }
        }
        f.out[ip] = val;
    }
}

}  // namespace gpvi
