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
// That holds for 2^-16 <= |b| <= 2^16 and a == 0 or 2^-96 <= |a| <= 2^96.
// The range test admits a first dividend in [2^-80, 2^80); its quotient, the
// dividend of the second division of (a/b)/b, then lies in [2^-96, 2^96] (or
// is 0), so one test of a covers both divisions.  The host checks the divisor
// range and verifies the identity over a full binade and at the ends of the
// range before it sets UnitDesc::fastdiv.  `hi` is 2^80 when the job may use
// the identity and 0 when it may not (then no value is in range).
//
// When any lane of a wave is out of range (in practice: dividends in the
// subnormal tail ahead of a wavefront), the whole wave divides as
// RN_f(RN_d(a * RN_d(1/b))), which equals the IEEE quotient RN_f(a/b) for
// every float a and b: the double product is within 2^-52 (relative) of
// a/b, while a/b, a quotient of two 24-bit significands, is either a float
// or at least 2^-49 (relative) away from the nearest float rounding
// boundary, so the two round alike.  (A CPU test compares the two over the
// subnormals and every significand of several binades.)
__device__ __forceinline__ bool cdiv_in_range(float a, float hi)
{
    const float aa = __builtin_fabsf(a);
    return (aa < hi) & ((aa >= 0x1p-80f) | (aa == 0.f));
}

__device__ __forceinline__ float div_hi(const UnitDesc& u) { return u.fastdiv ? 0x1p80f : 0.f; }

// Per-job constants of the point update, read from the unit descriptor once
// per work item.  (Read inside the row loop, every field would be fetched
// again after each store: the compiler cannot rule out that the buffer
// stores alias the descriptor.)
struct PointConsts {
    float dx, dz, dt, rdx, rdz, rdt, hi;
    double ydx, ydz, ydt;
    int run_fwi, k, ntx, ntz, pml, pitch4;
};

__device__ __forceinline__ PointConsts point_consts(const UnitDesc& u)
{
    return PointConsts{u.dx, u.dz, u.dt, u.rdx, u.rdz, u.rdt, div_hi(u), u.ydx, u.ydz, u.ydt, u.run_fwi, u.k,
                       u.ntx, u.ntz, u.pml, u.pitch * 4};
}

// RN_f(a/b) as RN_f(RN_d(a * yd)), yd = RN_d(1/b): exact for every float.
__device__ __forceinline__ float ddiv(float a, double yd) { return (float)((double)a * yd); }

__device__ __forceinline__ float cdiv_fast(float a, float b, float y)
{
    const float q0 = a * y;
    const float r = __builtin_fmaf(-q0, b, a);
    return __builtin_fmaf(r, y, q0);
}

// The fallback is taken by a wave-wide vote: when every active lane is in
// range (the common case) the quotient costs one scalar branch; otherwise
// the whole wave forms the exact double-product quotient.
__device__ __forceinline__ bool any_lane(bool x) { return __builtin_amdgcn_ballot_w64(x) != 0; }

// (a / b) / b, as the reference writes d2p/dx/dx and (.)/dt/dt.
__device__ __forceinline__ float div2_const(float a, float b, float y, float hi, double yd)
{
    const float q = cdiv_fast(cdiv_fast(a, b, y), b, y);
    const bool ok = cdiv_in_range(a, hi);
    if (__builtin_expect(!any_lane(!ok), 1)) return q;
    return ddiv(ddiv(a, yd), yd);
}

__device__ __forceinline__ float div1_const(float a, float b, float y, float hi, double yd)
{
    const float q = cdiv_fast(a, b, y);
    const bool ok = cdiv_in_range(a, hi);
    if (__builtin_expect(!any_lane(!ok), 1)) return q;
    return ddiv(a, yd);
}

// (a / bx) / bx + (c / bz) / bz with one range test for both quotients.
// Lanes that are not `active` (their result is discarded) do not vote.
__device__ __forceinline__ float lap_const(float a, float bx, float yx, double ydx, float c, float bz, float yz,
                                           double ydz, float hi, bool active = true)
{
    const float q = cdiv_fast(cdiv_fast(a, bx, yx), bx, yx) + cdiv_fast(cdiv_fast(c, bz, yz), bz, yz);
    const bool ok = (cdiv_in_range(a, hi) & cdiv_in_range(c, hi)) | !active;
    if (__builtin_expect(!any_lane(!ok), 1)) return q;
    return ddiv(ddiv(a, ydx), ydx) + ddiv(ddiv(c, ydz), ydz);
}

// Raw buffer access: a 128-bit scalar resource (base address, byte range,
// 32-bit data format) plus a per-lane byte offset (VGPR) and a uniform
// per-row byte offset (SGPR).  One instruction per access, no per-lane
// address arithmetic.  Bases sit at the first margin row of a field buffer
// (or the start of an inner-grid array) so every offset is non-negative.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int kRsrcFormat = 0x00020000;  // DATA_FORMAT = 32 bit (gfx9 family)

__device__ __forceinline__ Rsrc make_rsrc(const void* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, 0x7fffffff, kRsrcFormat);
}
// The b32 builtins move 32-bit words (unsigned int): bit-cast, never convert.
__device__ __forceinline__ float bload(Rsrc r, unsigned voff, int soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void bstore(float v, Rsrc r, unsigned voff, int soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 0);
}

// Order-2 transmitting extrapolation (fdtd2D_modelling.c:840-846): q1..q3
// are p1 at 0, 1, 2 cells inward, a0..a4 are p0 at 0..4 cells inward.
// abc_extrapolate reads a0..a4 through p0's resource at lane offset vo and
// row offsets ro + j*dstep (dstep: the byte step one cell inward, +-4 along
// x or +-pitch*4 along z; the margins keep every offset non-negative).
__device__ __forceinline__ float abc_extrap(float q1, float q2, float q3, float a0, float a1, float a2, float a3,
                                            float a4, float t11, float t12, float t13)
{
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

__device__ __forceinline__ float abc_extrapolate(float q1, float q2, float q3, Rsrc p0, unsigned vo, int ro,
                                                 int dstep, float t11, float t12, float t13)
{
    const float a0 = bload(p0, vo, ro), a1 = bload(p0, vo, ro + dstep), a2 = bload(p0, vo, ro + 2 * dstep);
    const float a3 = bload(p0, vo, ro + 3 * dstep), a4 = bload(p0, vo, ro + 4 * dstep);
    return abc_extrap(q1, q2, q3, a0, a1, a2, a3, a4, t11, t12, t13);
}

// w*H + (1.0 - w)*U: float product, double blend, float store (fdtd.c:840-847).
__device__ __forceinline__ float abc_blend(float w, float h, float u)
{
    return (float)((double)(w * h) + (1.0 - (double)w) * (double)u);
}

// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Field buffers of a unit's step: the pass's 4-buffer set (forward or
// adjoint) rotated by the pass-local step, with the zero field or (when a
// segment is recomputed) that segment's checkpoint pair as pre-history.
__device__ __forceinline__ FieldPtrs step_fields(const UnitDesc& u, const StepInfo& si)
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
    return rotate_ptrs(base, fe, pre0, pre1, si.q);
}

// One time step of a macro, as seen by the points a wave updates.
//
// Imaging in pairs.  The reference adds each imaging step's term to the
// gradient as it goes (fdtd2D_modelling.c:959-975), one read-modify-write of
// the gradient per imaging step.  Here, for nt_interval k <= 2, the terms of
// two consecutive imaging steps it + k and it of one segment (local slots
// 2m + 1 and 2m) are applied at step it: the field of step it + k is still in
// this step's inputs (p1 for k = 1, p0 for k = 2) and its store slot is next
// to this one's.  The two terms are added one after the other, each rounded
// as the reference rounds it, so the gradient is the same bits; the step at
// the odd slot skips imaging, and the gradient is read and written once per
// pair.  A segment with an odd number of slots images its top slot alone.
struct StepCtx {
    int it, mode;
    bool first_image;  // first gradient write of the adjoint pass: the gradient starts from 0
    bool has_save;     // wavefield store (forward-like) or imaging (adjoint) at this step
    bool img_pair;     // imaging: also apply the term of step it + k (store slot + 1)
    bool has_ck;       // checkpoint (p0, p1) written at this step
    int slot4, fe4;    // byte distance to the next store slot (save + slot4: slot + 1) and to ck1 (ck0 + fe4)
    Rsrc out, save, ck0;
};

__device__ __forceinline__ StepCtx make_step(const UnitDesc& u, int mode, int it, int seg, float* out,
                                             int64_t margin)
{
    StepCtx c;
    c.it = it;
    c.mode = mode;
    const bool fwd_like = mode != kAdj;
    const bool on_k = (it % u.k) == 0;
    const bool store = fwd_like && u.store && on_k && (mode == kRec || u.nseg == 1);
    const int slot = (u.nseg == 1) ? it / u.k : (it - seg * u.ckpt) / u.k;   // store slot in the segment
    // store slots of this segment
    const int seg_len = (u.nseg == 1) ? u.nt : (seg == u.nseg - 1 ? u.last_len : u.ckpt);
    const int nslots = (seg_len + u.k - 1) / u.k;
    const bool pairs = u.k <= 2;
    const bool image = !fwd_like && on_k && !(pairs && (slot & 1));   // odd slots: applied with slot - 1
    c.has_save = store || image;
    c.img_pair = image && pairs && slot + 1 < nslots;
    if (c.has_save) c.save = make_rsrc(u.save + (int64_t)slot * u.nx * u.nz);
    c.slot4 = u.nx * u.nz * 4;
    c.fe4 = (int)(u.field_elems * 4);
    c.has_ck = (mode == kFwd) && (u.nseg > 1) && (it > 0) && (it % u.ckpt == 0);
    if (c.has_ck) {
        const int j = it / u.ckpt;
        c.ck0 = make_rsrc(u.ck + (int64_t)(2 * j) * u.field_elems - margin);   // ck1 follows at fe4 bytes
    }
    // the first gradient write: the highest imaging step that images, in the top segment
    const int first_slot = pairs ? ((nslots - 1) & ~1) : nslots - 1;
    c.first_image = u.grad_init && image && seg == u.nseg - 1 && slot == first_slot;
    c.out = make_rsrc(out - margin);
    return c;
}

// The value of one point after the stencil, the ABC and the source or
// residual injection.  zn[H + m] is p1 at row iz + m (|m| <= H, this lane's
// column), sp[m] p1 at column ix + m of row iz (the LDS row).  INSIDE: every
// lane of the wave is an inner-grid point away from the absorbing bands.
// rp0/vo/ro: p0 of this step's resource with this lane's and this row's
// byte offsets (inward reads of the ABC).
template <int LC, bool INSIDE, bool BAND = false>
__device__ __forceinline__ float point_value(const UnitDesc& u, const PointConsts& kc, const StepCtx& c, int ix, int iz,
                                             const float (&zn)[2 * LC - 1], const float* sp, float p0c, float v,
                                             bool src_row, bool rec_row, Rsrc rp0, unsigned vo, int ro,
                                             bool active = true, const float* s0 = nullptr, float wl = 0.f,
                                             float wr = 0.f)
{
    constexpr int H = LC - 1;
    const float hi = kc.hi;
    const float p1c = zn[H];
    const int ntx = kc.ntx, ntz = kc.ntz, pml = kc.pml;
    float val = p1c;
    if (INSIDE || (iz >= LC && iz < ntz - LC && ix >= LC && ix < ntx - LC)) {
        // S1: (2*p1 - p0) + (v*dt)^2 * (d2x/dx/dx + d2z/dz/dz)
        float dxx = fd_weight(LC, 0) * p1c;
#pragma unroll
        for (int m = 1; m < LC; ++m) dxx = dxx + fd_weight(LC, m) * (sp[m] + sp[-m]);
        float dzz = fd_weight(LC, 0) * p1c;
#pragma unroll
        for (int m = 1; m < LC; ++m) dzz = dzz + fd_weight(LC, m) * (zn[H + m] + zn[H - m]);
        const float vdt = v * kc.dt;
        const float lap = lap_const(dxx, kc.dx, kc.rdx, kc.ydx, dzz, kc.dz, kc.rdz, kc.ydz, hi, active);
        val = (2.f * p1c - p0c) + (vdt * vdt) * lap;
    }
    if (!INSIDE) {
        // S2: top, bottom, left, right, in that order.
        const bool top = iz >= LC && iz < pml;
        const bool bot = iz >= ntz - pml && iz < ntz - LC;
        const bool lef = ix >= LC && ix < pml;
        const bool rig = ix >= ntx - pml && ix < ntx - LC;
        if (top || bot || lef || rig) {
            const int pstep = kc.pitch4;
            const float sc = div1_const(v * kc.dt, kc.dx, kc.rdx, hi, kc.ydx);
            const float t11 = ((2.f - sc) * (1.f - sc)) / 2.f;
            const float t12 = sc * (2.f - sc);
            const float t13 = (sc * (sc - 1.f)) / 2.f;
            const Rsrc rw = make_rsrc(u.wtab);
            if (top)
                val = abc_blend(bload(rw, (unsigned)(iz - LC) * 4u, 0),
                                abc_extrapolate(p1c, zn[H + 1], zn[H + 2], rp0, vo, ro, pstep, t11, t12, t13), val);
            if (bot)
                val = abc_blend(bload(rw, (unsigned)(ntz - LC - iz - 1) * 4u, 0),
                                abc_extrapolate(p1c, zn[H - 1], zn[H - 2], rp0, vo, ro, -pstep, t11, t12, t13), val);
            if (BAND) {
                // Side-band items: p0 of the lane's row is staged in LDS with
                // 4 columns on each side (s0 = this lane's column) and the
                // lane's x taper weights are read once per item.
                if (lef)
                    val = abc_blend(wl, abc_extrap(p1c, sp[1], sp[2], s0[0], s0[1], s0[2], s0[3], s0[4], t11, t12, t13),
                                    val);
                if (rig)
                    val = abc_blend(wr, abc_extrap(p1c, sp[-1], sp[-2], s0[0], s0[-1], s0[-2], s0[-3], s0[-4], t11, t12,
                                                   t13),
                                    val);
            } else {
                if (lef)
                    val = abc_blend(bload(rw, (unsigned)(ix - LC) * 4u, 0),
                                    abc_extrapolate(p1c, sp[1], sp[2], rp0, vo, ro, 4, t11, t12, t13), val);
                if (rig)
                    val = abc_blend(bload(rw, (unsigned)(ntx - LC - ix - 1) * 4u, 0),
                                    abc_extrapolate(p1c, sp[-1], sp[-2], rp0, vo, ro, -4, t11, t12, t13), val);
            }
        }
    }
    const float dt = kc.dt;
    if (c.mode != kAdj) {
        // S3: source after the ABC.
        if (src_row && ix == u.s_ix) val = val + G(u.wav)[c.it] * dt * dt * v * v;
    } else if (rec_row) {
        // S5: residual injection at the receivers.
        const int d = ix - u.r_ix0;
        if (d >= 0 && d % u.dr == 0 && d / u.dr < u.nr) {
            const int ri = c.it * u.nr + d / u.dr;
            const float res = kc.run_fwi ? G(u.syn)[ri] - G(u.obs)[ri] : G(u.obs)[ri];
            val = val + res * dt * dt * v * v;
        }
    }
    return val;
}

// Outputs of one owned point: checkpoint, wavefield store and record
// (forward-like steps) or imaging condition (adjoint steps), and the field.
// vo/ro: lane/row byte offsets in padded field rows; gvo/go: lane/row byte
// offsets in the inner-grid arrays (used when `inner`: the point lies in the
// inner grid).
__device__ __forceinline__ void store_point(const UnitDesc& u, const PointConsts& kc, const StepCtx& c, float val,
                                            float p0c, float p1c, float v, int ix, unsigned vo, int ro,
                                            unsigned gvo, int go, Rsrc grad, bool rec_row, bool inner)
{
    const float dt = kc.dt;
    if (c.mode != kAdj) {
        if (c.has_ck) {
            bstore(p0c, c.ck0, vo, ro);
            bstore(p1c, c.ck0, vo, ro + c.fe4);
        }
        if (c.has_save && inner) {
            float sv = val;
            if (kc.run_fwi) sv = div2_const((val - 2.f * p1c) + p0c, dt, kc.rdt, kc.hi, kc.ydt);
            bstore(sv, c.save, gvo, go);
        }
        if (c.mode == kFwd && rec_row) {
            const int d = ix - u.r_ix0;
            if (d >= 0 && d % u.dr == 0 && d / u.dr < u.nr) G(u.syn)[c.it * u.nr + d / u.dr] = val;
        }
    } else if (c.has_save && inner) {
        // imaging: the term of step it + k first (when paired), then step it's
        const float sv = bload(c.save, gvo, go);
        const float sv2 = c.img_pair ? bload(c.save, gvo, go + c.slot4) : 0.f;
        const float lam2 = kc.k == 1 ? p1c : p0c;   // the field of step it + k
        float g = c.first_image ? 0.f : bload(grad, gvo, go);
        if (kc.run_fwi) {
            const double vd = (double)v;
            const float v3 = (float)(vd * vd * vd);
            const double coef = -2.0 / (double)v3;
            if (c.img_pair) g = (float)((double)g + coef * (double)sv2 * (double)lam2 * (double)kc.k * (double)dt);
            g = (float)((double)g + coef * (double)sv * (double)val * (double)kc.k * (double)dt);
        } else {
            if (c.img_pair) g = g + sv2 * lam2 * (float)kc.k * dt;
            g = g + sv * val * (float)kc.k * dt;
        }
        bstore(g, grad, gvo, go);
    }
    bstore(val, c.out, vo, ro);
}

// LDS floats one wavefront uses: the rows of a row strip (64 + 2H each;
// kRowBlock rows for a single step, whose interior blocks stage a block at
// once, TB rows for a pair) or the RS row slots of a side-band item (p1 in
// W + 2H and p0 in W + 8 floats each, RS*W = 64, RS <= 4).
__host__ __device__ constexpr int wave_lds_floats(int lc, int tb)
{
    return (tb == 1 ? kRowBlock : tb) * (64 + 2 * (lc - 1)) > 128 + 4 * (2 * (lc - 1) + 8)
               ? (tb == 1 ? kRowBlock : tb) * (64 + 2 * (lc - 1))
               : 128 + 4 * (2 * (lc - 1) + 8);
}

// One side-band item (one time step): columns [0, W) (item & 1 == 0) or
// [ntx - W, ntx) of rows [zb, ze), W = 64/RS >= pml (that side's width in
// UnitDesc::band_w).
// Lane l owns column l % W of row zb + RS*g + l / W at wave step g, so every
// wave step advances RS rows and most lanes are absorbing-band points.  Each
// lane keeps p1 of its own column around its row in registers (z
// neighbours); the RS rows of a step are staged in LDS with their halo
// columns (x neighbours).  Offsets: a lane's row varies across the wave, so
// every access is a per-lane byte offset from the lane's reference row
// rb = zb + l/W - H - 4 plus a uniform multiple of the pitch (>= 0).
template <int LC, int RS>
__device__ __forceinline__ void process_band(const UnitDesc& u, const PointConsts& kc, const StepCtx& c,
                                             const FieldPtrs& f, int item, float* lds)
{
    constexpr int H = LC - 1;
    constexpr int W = 64 / RS;
    constexpr int RWb = W + 2 * H;
    const int side = item & 1;
    const int zb = (item >> 1) * u.zseg;
    const int ze = min(zb + u.zseg, kc.ntz);
    const int ntx = kc.ntx, ntz = kc.ntz, pml = kc.pml, pitch4 = kc.pitch4;
    const int lane = threadIdx.x & 63;
    const int cl = lane % W, j = lane / W;
    const int col0 = side ? ntx - W : 0;
    const int ix = col0 + cl;
    const int64_t margin = (int64_t)kPadRows * u.pitch;
    const Rsrc rp1 = make_rsrc(f.p1 - margin);
    const Rsrc rp0 = make_rsrc(f.p0 - margin);
    const Rsrc rv = make_rsrc(u.vel - margin);
    const Rsrc rg = make_rsrc(u.grad);
    const int rb = zb + j - H - 4;
    const unsigned lb = (unsigned)((rb + kPadRows) * pitch4 + ix * 4);
    const bool halo_lane = cl < H;
    const unsigned hslot = side ? cl : W + H + cl;
    const int hcol = side ? ntx - W - H + cl : W + cl;
    const unsigned hb = (unsigned)((rb + kPadRows) * pitch4 + hcol * 4);
    const int ro_abc = 4 * pitch4 + 16;   // inward ABC reads stay at offsets >= 0
    const int r_iz = u.r_iz, s_iz = u.s_iz, nx = u.nx;
    const bool inner_x = ix >= pml && ix < ntx - pml;
    float* myrow = lds + j * RWb;
    const unsigned li = cl + H;
    // p0 of the lane's row in LDS, columns col0 - 4 .. col0 + W + 3 (the x
    // extrapolation reads up to 4 cells inward); lanes cl < 8 also load the
    // 4 columns on each side.
    constexpr int RW0 = W + 8;
    float* myrow0 = lds + RS * RWb + j * RW0;
    const bool halo0_lane = cl < 8;
    const unsigned hslot0 = cl < 4 ? cl : W + cl;
    const unsigned hb0 = (unsigned)((rb + kPadRows) * pitch4 + (col0 - 4 + (int)hslot0) * 4);
    // x taper weights of the lane's column (fdtd.c:870-900), fixed for the item
    const Rsrc rw = make_rsrc(u.wtab);
    const float wl = (ix >= LC && ix < pml) ? bload(rw, (unsigned)(ix - LC) * 4u, 0) : 0.f;
    const float wr = (ix >= ntx - pml && ix < ntx - LC) ? bload(rw, (unsigned)(ntx - LC - ix - 1) * 4u, 0) : 0.f;

    // w[m]: p1 at row (lane's row) - H + m; the lane's row at step g is rb + H + 4 + RS*g
    float w[2 * H + 1];
#pragma unroll
    for (int m = 0; m <= 2 * H; ++m) w[m] = bload(rp1, lb, (4 + m) * pitch4);
    float hal = halo_lane ? bload(rp1, hb, (H + 4) * pitch4) : 0.f;
    float p0c = bload(rp0, lb, (H + 4) * pitch4);
    float hal0 = halo0_lane ? bload(rp0, hb0, (H + 4) * pitch4) : 0.f;
    float vc = bload(rv, lb, (H + 4) * pitch4);
    const int ng = (ze - zb + RS - 1) / RS;
    for (int g = 0; g < ng; ++g) {
        const int base = RS * g;
        const bool more = g + 1 < ng;
        float nw[RS], nh = 0.f, n0 = 0.f, nh0 = 0.f, nv = 0.f;
        if (more) {
#pragma unroll
            for (int k = 0; k < RS; ++k) nw[k] = bload(rp1, lb, (base + 2 * H + 5 + k) * pitch4);
            nh = halo_lane ? bload(rp1, hb, (base + RS + H + 4) * pitch4) : 0.f;
            n0 = bload(rp0, lb, (base + RS + H + 4) * pitch4);
            nh0 = halo0_lane ? bload(rp0, hb0, (base + RS + H + 4) * pitch4) : 0.f;
            nv = bload(rv, lb, (base + RS + H + 4) * pitch4);
        } else {
#pragma unroll
            for (int k = 0; k < RS; ++k) nw[k] = 0.f;
        }
        const int iz = zb + base + j;
        myrow[li] = w[H];
        if (halo_lane) myrow[hslot] = hal;
        myrow0[cl + 4] = p0c;
        if (halo0_lane) myrow0[hslot0] = hal0;
        __builtin_amdgcn_wave_barrier();
        if (iz < ze) {
            const bool rec_row = iz == r_iz, src_row = iz == s_iz;
            const unsigned vo = lb + (unsigned)((base + H + 4) * pitch4 - ro_abc);
            const float val = point_value<LC, false, true>(u, kc, c, ix, iz, w, myrow + li, p0c, vc, src_row, rec_row,
                                                           rp0, vo, ro_abc, true, myrow0 + cl + 4, wl, wr);
            const bool inner = inner_x && iz >= pml && iz < ntz - pml;
            const unsigned gv = inner ? (unsigned)(((iz - pml) * nx + (ix - pml)) * 4) : 0u;
            store_point(u, kc, c, val, p0c, w[H], vc, ix, vo, ro_abc, gv, 0, rg, rec_row, inner);
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int m = 0; m + RS <= 2 * H; ++m) w[m] = w[m + RS];
#pragma unroll
        for (int k = 0; k < RS; ++k) w[2 * H + 1 - RS + k] = nw[k];
        hal = nh;
        p0c = n0;
        hal0 = nh0;
        vc = nv;
    }
}

// One wavefront: one strip segment of one unit at local macro step s.
//
// tb == 1: lane l updates column x0 + l, rows [zb, ze), one step.
// tb == 2: lane l covers column x0 - H + l.  Step A (the macro's first
//   step) runs on every lane for rows [zb - H, ze + H); step B (the second)
//   runs H rows behind it on the same lanes, reading A's results from a
//   second register window (z-neighbours) and a second LDS row
//   (x-neighbours).  Lanes H .. 63-H (the strip's 64 - 2H owned columns)
//   store; A's values on the other lanes and rows only feed B, and are
//   recomputed bit-identically by the neighbouring strips and segments.
//
// TB (1 or 2) is the unit's steps per macro step, UnitDesc::tb.  `lds` is
// the wave's private scratch: rows of 64 + 2H floats (wave_lds_floats).
template <int LC, int TB>
__device__ __forceinline__ void process_item(const UnitDesc& u, int s, int item, float* lds)
{
    constexpr int H = LC - 1;
    constexpr int R = kRowBlock;
    constexpr int NW = 2 * H + R;      // rows of a register window per block
    constexpr int RW = 64 + 2 * H;     // LDS row length
    const StepInfo si = decode_step(u, s);
    if (si.mode == kIdle) return;
    const bool pair = TB == 2 && si.count == 2;

    const int ntx = u.ntx, ntz = u.ntz, pitch = u.pitch, pml = u.pml;
    const int64_t fe = u.field_elems;
    const int64_t margin = (int64_t)kPadRows * pitch;
    float* base = si.mode == kAdj ? u.adj : u.fwd;
    const FieldPtrs f = step_fields(u, si);
    const StepCtx cA = make_step(u, si.mode, si.it, si.seg, f.out, margin);
    const int itB = si.mode == kAdj ? si.it - 1 : si.it + 1;
    StepCtx cB;
    if (pair) cB = make_step(u, si.mode, itB, si.seg, base + fe * ((si.q + 1) & 3), margin);

    // Side-band items first (UnitDesc::band_w = left width | right width << 8),
    // then row strips over [left width, ntx - right width).
    const int bl = TB == 1 ? (u.band_w & 0xff) : 0;
    const int br = TB == 1 ? (u.band_w >> 8) : 0;
    if (bl > 0) {
        if (item < 2 * u.band_tz) {
            if ((item & 1 ? br : bl) == 16)
                process_band<LC, 4>(u, point_consts(u), cA, f, item, lds);
            else
                process_band<LC, 2>(u, point_consts(u), cA, f, item, lds);
            return;
        }
        item -= 2 * u.band_tz;
    }
    const int seg = item / u.tiles_x;
    const int x0 = bl + (item - seg * u.tiles_x) * u.stripw;
    const int zb = seg * u.zseg;
    const int ze = min(zb + u.zseg, ntz);
    const int lane = threadIdx.x & 63;
    const int loff = TB == 2 ? H : 0;             // first owned lane
    const int ix = x0 - loff + lane;              // this lane's column
    const int x_end = ntx - br;                   // row strips own columns [bl, ntx - br)
    const bool lane_ok = ix >= 0 && ix < x_end;
    const bool owned = lane_ok && lane >= loff && lane < loff + u.stripw;
    const int cx0 = x0 - loff - H;                // column of LDS slot 0
    const unsigned li = lane + H;                 // this lane's LDS slot
    const unsigned vo = li * 4u;
    const bool halo_lane = lane < 2 * H;
    const unsigned hslot = lane < H ? lane : 64 + lane;
    // Lanes that are not halo lanes read the first halo lane's word (same
    // cache line) so the halo loads need no execution mask.
    const unsigned ho = halo_lane ? hslot * 4u : 0u;
    const int pitch4 = pitch * 4;
    auto row_off = [&](int z) { return (z + kPadRows) * pitch4 + cx0 * 4; };
    // Every owned lane of the wave inner and away from the bands (for both
    // steps); lanes past x_end compute values nobody stores.
    const bool strip_inside = (x0 - loff) >= pml && min(x0 - loff + 64, x_end) <= ntx - pml;
    const bool inner_x = ix >= pml && ix < ntx - pml;
    const unsigned gvo = (unsigned)(ix - pml) * 4u;   // lane offset in inner-grid arrays (inner_x lanes)
    const PointConsts kc = point_consts(u);
    const int r_iz = u.r_iz, s_iz = u.s_iz, nx4 = u.nx * 4;

    const Rsrc rp1 = make_rsrc(f.p1 - margin);
    const Rsrc rp0 = make_rsrc(f.p0 - margin);
    const Rsrc rv = make_rsrc(u.vel - margin);
    const Rsrc rg = make_rsrc(u.grad);
    float* row1 = lds;
    float* row2 = lds + (TB - 1) * RW;

    // Rows of step A: the segment, plus H on each side for a pair.
    const int za_begin = pair ? zb - H : zb;
    const int za_end = pair ? ze + H : ze;

    float w1[NW];      // p1 at rows za0-H .. za0+R+H-1 (step A's z-neighbours)
    float w2[NW];      // A's result at rows za0-2H .. za0+R-1 (step B's z-neighbours)
    float vh[H + R];   // velocity at rows za0-H .. za0+R-1
    float hal[R], p0r[R];
#pragma unroll
    for (int j = 0; j < NW; ++j) {
        w1[j] = bload(rp1, vo, row_off(za_begin - H + j));
        w2[j] = 0.f;
    }
#pragma unroll
    for (int j = 0; j < H; ++j) vh[j] = 0.f;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int ro = row_off(za_begin + r);
        hal[r] = bload(rp1, ho, ro);
        p0r[r] = bload(rp0, vo, ro);
        vh[H + r] = bload(rv, vo, ro);
    }

    for (int za0 = za_begin; za0 < za_end; za0 += R) {
        // Issue every load of the next block (if there is one) before
        // computing this one.
        const int za1 = za0 + R;
        float n1[R], nh[R], n0[R], nv[R];
        if (za1 < za_end) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int ro = row_off(za1 + r);
                n1[r] = bload(rp1, vo, row_off(za1 + H + r));
                nh[r] = bload(rp1, ho, ro);
                n0[r] = bload(rp0, vo, ro);
                nv[r] = bload(rv, vo, ro);
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) n1[r] = nh[r] = n0[r] = nv[r] = 0.f;
        }

        // A block of R interior rows (single step, every owned lane inner,
        // no source or receiver row) takes the lean path: its R rows are
        // staged in LDS at once and updated with no per-row tests.  Imaging
        // steps keep the per-row path, where each row's loads of the stored
        // field and the gradient are issued ahead of that row's arithmetic.
        if (TB == 1 && strip_inside && !(cA.has_save && cA.mode == kAdj) && za0 >= pml && za0 + R <= za_end &&
            za0 + R <= ntz - pml && (unsigned)(r_iz - za0) >= (unsigned)R && (unsigned)(s_iz - za0) >= (unsigned)R) {
#pragma unroll
            for (int r = 0; r < R; ++r) lds[r * RW + li] = w1[r + H];
            if (halo_lane) {
#pragma unroll
                for (int r = 0; r < R; ++r) lds[r * RW + hslot] = hal[r];
            }
            __builtin_amdgcn_wave_barrier();
            float val[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float zA[2 * H + 1];
#pragma unroll
                for (int m = 0; m <= 2 * H; ++m) zA[m] = w1[r + m];
                val[r] = point_value<LC, true>(u, kc, cA, ix, za0 + r, zA, lds + r * RW + li, p0r[r], vh[H + r], false,
                                               false, rp0, vo, row_off(za0 + r), lane_ok);
            }
            __builtin_amdgcn_wave_barrier();
            if (owned) {
                if (!cA.has_save && !cA.has_ck) {
                    // a plain step (no store, checkpoint or imaging): the field only
#pragma unroll
                    for (int r = 0; r < R; ++r) bstore(val[r], cA.out, vo, row_off(za0 + r));
                } else {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        store_point(u, kc, cA, val[r], p0r[r], w1[r + H], vh[H + r], ix, vo, row_off(za0 + r), gvo,
                                    (za0 + r - pml) * nx4, rg, false, inner_x);
                }
            }
        } else
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int za = za0 + r;
            if (za >= za_end) break;
            // ---- step A at row za (every lane) ----
            float zA[2 * H + 1];
#pragma unroll
            for (int m = 0; m <= 2 * H; ++m) zA[m] = w1[r + m];
            row1[li] = zA[H];
            if (halo_lane) row1[hslot] = hal[r];
            __builtin_amdgcn_wave_barrier();
            float valA = 0.f;
            if (za >= 0 && za < ntz) {
                const bool row_inner = za >= pml && za < ntz - pml;
                const bool rec_row = za == r_iz;
                const bool src_row = za == s_iz;
                const float* sp = row1 + li;
                if (strip_inside && row_inner) {
                    valA = point_value<LC, true>(u, kc, cA, ix, za, zA, sp, p0r[r], vh[H + r], src_row, rec_row, rp0, vo,
                                                  row_off(za), lane_ok);
                } else if (lane_ok) {
                    valA = point_value<LC, false>(u, kc, cA, ix, za, zA, sp, p0r[r], vh[H + r], src_row, rec_row, rp0, vo,
                                                   row_off(za));
                }
                if (owned && za >= zb && za < ze)
                    store_point(u, kc, cA, valA, p0r[r], zA[H], vh[H + r], ix, vo, row_off(za), gvo, (za - pml) * nx4,
                                rg, rec_row, row_inner && inner_x);
            }
            w2[2 * H + r] = valA;
            __builtin_amdgcn_wave_barrier();
            // ---- step B at row zbr = za - H ----
            const int zbr = za - H;
            if (pair && zbr >= zb) {
                float zB[2 * H + 1];
#pragma unroll
                for (int m = 0; m <= 2 * H; ++m) zB[m] = w2[r + m];
                row2[li] = zB[H];
                __builtin_amdgcn_wave_barrier();
                const bool row_inner = zbr >= pml && zbr < ntz - pml;
                const bool rec_row = zbr == r_iz;
                const bool src_row = zbr == s_iz;
                const float* sp = row2 + li;
                const float p0b = w1[r];   // step A's p1 at row zbr
                const float vb = vh[r];
                float valB = 0.f;
                if (strip_inside && row_inner) {
                    valB = point_value<LC, true>(u, kc, cB, ix, zbr, zB, sp, p0b, vb, src_row, rec_row, rp1, vo, row_off(zbr));
                } else if (lane_ok) {
                    valB = point_value<LC, false>(u, kc, cB, ix, zbr, zB, sp, p0b, vb, src_row, rec_row, rp1, vo,
                                                   row_off(zbr));
                }
                if (owned)
                    store_point(u, kc, cB, valB, p0b, zB[H], vb, ix, vo, row_off(zbr), gvo, (zbr - pml) * nx4, rg,
                                rec_row, row_inner && inner_x);
                __builtin_amdgcn_wave_barrier();
            }
        }

        // Slide the windows to the next block.
#pragma unroll
        for (int j = 0; j < 2 * H; ++j) {
            w1[j] = w1[j + R];
            w2[j] = w2[j + R];
        }
#pragma unroll
        for (int j = 0; j < H; ++j) vh[j] = vh[j + R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            w1[2 * H + r] = n1[r];
            w2[2 * H + r] = 0.f;
            hal[r] = nh[r];
            p0r[r] = n0[r];
            vh[H + r] = nv[r];
        }
    }
}

}  // namespace gpvi
