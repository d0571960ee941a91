// Device-side data layout and the fused time-step of the 2-D acoustic
// FDTD misfit/gradient solver (gfx950).
//
// One "unit" is one (velocity sample, shot) pair: its own four rotating
// forward fields, four adjoint fields, checkpoints, wavefield store and
// private gradient, all in HBM.  Every launch advances a set of units by one
// macro step (one or two time steps) each; every lane owns a grid column and
// does, for each of its points, the whole per-step work of the reference
// solver:
//
//   interior Laplacian update   fdtd2D_modelling.c:756-780
//   4 ordered ABC sweeps        fdtd2D_modelling.c:833-900
//   source / residual injection fdtd2D_modelling.c:910-911, 949-957
//   wavefield store, records    fdtd2D_modelling.c:913-937
//   imaging condition           fdtd2D_modelling.c:959-975
//
// The copy rotation of the reference (p0<-p1<-p2) becomes a rotation of
// buffer indices, with the reference's invariant p2 == p1 at the start of a
// step turned into "a point the stencil does not update starts from p1".
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpvi {

constexpr int kMaxLc = 7;       // solver half-width (configured Lc 3..6, plus one)
constexpr int kPadRows = 16;    // margin rows above and below every field buffer
constexpr int kRowBlock = 4;    // rows per prefetch block of a strip segment

// Central second-derivative weights for solver half-width lc = 4..7
// (configured Lc 3..6), weight m = 0..lc-1: the exact rationals of the
// reference's table (fdtd2D_modelling.c:1272-1326) rounded to float.  One
// definition for the host (UnitDesc::fd, validation) and the kernels (where
// lc is a template parameter, so every weight is a literal).
__host__ __device__ constexpr float fd_weight(int lc, int m)
{
    switch (lc * 8 + m) {
    case 32: return (float)(-49.0 / 18); case 33: return (float)(3.0 / 2);
    case 34: return (float)(-3.0 / 20);  case 35: return (float)(1.0 / 90);
    case 40: return (float)(-205.0 / 72); case 41: return (float)(8.0 / 5);
    case 42: return (float)(-1.0 / 5);    case 43: return (float)(8.0 / 315);
    case 44: return (float)(-1.0 / 560);
    case 48: return (float)(-5269.0 / 1800); case 49: return (float)(5.0 / 3);
    case 50: return (float)(-5.0 / 21);      case 51: return (float)(5.0 / 126);
    case 52: return (float)(-5.0 / 1008);    case 53: return (float)(1.0 / 3150);
    case 56: return (float)(-5369.0 / 1800); case 57: return (float)(12.0 / 7);
    case 58: return (float)(-15.0 / 56);     case 59: return (float)(10.0 / 189);
    case 60: return (float)(-1.0 / 112);     case 61: return (float)(2.0 / 1925);
    case 62: return (float)(-1.0 / 16632);
    default: return 0.f;
    }
}

// Per-unit modes of one local step.
enum StepMode : int { kFwd = 0, kRec = 1, kAdj = 2, kIdle = 3 };

// Device unit descriptor.  Field buffers are "pitched": ntz rows of `pitch`
// floats (pitch >= ntx, a multiple of 64 floats); each row may start a few
// floats into its pitch, so that the row strips' first column is 128-byte
// aligned.
struct UnitDesc {
    int ntx, ntz, pitch, nx;
    int nz, pml, pml0, lc;
    int nt, k, nr, dr;
    int r_ix0, r_iz, s_ix, s_iz;
    int nseg, ckpt, last_len, nslot;   // segments of the adjoint schedule
    int tiles_x, tiles_z, start, s_off;  // tiles_x: column strips, tiles_z: row segments;
                                         // local macro step s = global - start + s_off
    int zseg, tb, stripw, band_w;        // rows per strip segment (one wavefront's work item),
                                         // time steps per macro (1 or 2), columns per strip,
                                         // width of the side-band items (0: none)
    int total, run_fwi, grad_init, store;  // total macro steps; grad_init: first imaging step writes;
                                           // store: keep wavefields
    float dt, dx, dz, rdt;                 // rdx/rdz/rdt = RN(1/dx) ...: exact constant division
    float rdx, rdz;
    int fastdiv, band_tz;                  // 1: the host verified the reciprocal division for dx, dz, dt;
                                           // side-band items per side
    float fd[8];
    double ydx, ydz, ydt;  // RN(1/dx), RN(1/dz), RN(1/dt) in double: the exact fallback division
    int64_t field_elems;  // stride between field buffers: (ntz + 2*kPadRows) * pitch
    const float* wtab;    // ABC taper w[i] = 1 - i/pml0, i < pml0
    // Field pointers point at row 0, column 0 of a buffer that has kPadRows
    // rows of margin before and after, so every row a strip segment touches is
    // addressable without bounds checks (margin values never reach a grid
    // point of the solution).
    float* fwd;           // 4 field buffers
    float* adj;           // 4 field buffers
    float* ck;            // nseg * 2 field buffers (segment 0 slot unused)
    float* save;          // nslot * nx * nz, compact [slot][iz][ix] inner grid
    float* grad;          // nx * nz
    const float* vel;     // one field buffer, padded by edge replication
    const float* zero;    // one all-zero field buffer (the fields before step 0)
    const float* wav;     // nt
    float* syn;           // nt * nr
    const float* obs;     // nt * nr
};

// Global step counter and tile count of the multi-unit launches; lives in
// device memory so a captured graph segment can be replayed unchanged.
// Tiles are grouped by solver half-width (lc = 4..7): group g = lc - 4 holds
// ntiles_lc[g] tiles starting at tile_off[g] of the tile list.
struct StepCtl {
    int step;
    int pad[3];
    int ntiles_lc[4];
    int tile_off[4];
};

struct TileRef {
    int unit;
    int tile;
};

// ---------------------------------------------------------------------------
// Schedule.  A unit advances by one "macro step" per launch: tb (1 or 2)
// consecutive time steps of one pass.  Passes, in order:
//   forward   it = 0 .. nt-1
//   then segments j = nseg-1 .. 0, each:
//     (nseg > 1) recompute it = jC .. jC+Lj-1 from checkpoint j
//     adjoint   it = jC+Lj-1 .. jC
// A pass of odd length ends with a single-step macro.  q is the pass-local
// step index of the first step (used for the buffer rotation); `count` is
// the number of steps (1 or tb) of the macro.
struct StepInfo {
    int mode, it, q, seg, count;
};

__host__ __device__ inline int cdiv_i(int a, int b) { return (a + b - 1) / b; }
__host__ __device__ inline int min_i(int a, int b) { return a < b ? a : b; }

// Number of macro steps of a unit (the launches it takes).
__host__ __device__ inline int macro_steps(int nt, int nseg, int ckpt, int last_len, int tb, bool forward,
                                           bool backward)
{
    int n = forward ? cdiv_i(nt, tb) : 0;
    if (backward)
        n += (nseg == 1) ? cdiv_i(nt, tb) : (nseg - 1) * 2 * cdiv_i(ckpt, tb) + 2 * cdiv_i(last_len, tb);
    return n;
}

__host__ __device__ inline StepInfo decode_step(const UnitDesc& u, int s)
{
    StepInfo r{kIdle, 0, 0, 0, 0};
    const int tb = u.tb;
    if (s < 0) return r;
    const int nf = cdiv_i(u.nt, tb);
    if (s < nf) {
        const int it = tb * s;
        r.mode = kFwd; r.it = it; r.q = it; r.seg = (u.nseg > 1) ? it / u.ckpt : 0;
        r.count = min_i(tb, u.nt - it);
        return r;
    }
    const int b = s - nf;
    if (u.nseg == 1) {
        if (b >= nf) return r;
        const int it = u.nt - 1 - tb * b;
        r.mode = kAdj; r.it = it; r.q = u.nt - 1 - it; r.seg = 0; r.count = min_i(tb, u.nt - tb * b);
        return r;
    }
    const int bl = 2 * cdiv_i(u.last_len, tb), bc = 2 * cdiv_i(u.ckpt, tb);
    int j, w, lj;
    if (b < bl) {
        j = u.nseg - 1; w = b; lj = u.last_len;
    } else {
        const int b2 = b - bl;
        j = u.nseg - 2 - b2 / bc;
        if (j < 0) return r;
        w = b2 % bc; lj = u.ckpt;
    }
    const int h = cdiv_i(lj, tb);
    r.seg = j;
    if (w < h) {
        r.mode = kRec; r.it = j * u.ckpt + tb * w; r.q = tb * w; r.count = min_i(tb, lj - tb * w);
    } else {
        const int w2 = w - h;
        r.mode = kAdj; r.it = j * u.ckpt + lj - 1 - tb * w2; r.q = u.nt - 1 - r.it;
        r.count = min_i(tb, lj - tb * w2);
    }
    return r;
}

// Four-buffer rotation with a two-field pre-history (zero, or a checkpoint):
//   q = 0: p0 = pre0, p1 = pre1, out = B0
//   q = 1: p0 = pre1, p1 = B0,   out = B1
//   q >= 2: p0 = B[(q-2)%4], p1 = B[(q-1)%4], out = B[q%4]
// With four buffers the second step of a macro (output B[(q+1)%4]) never
// writes a buffer the macro's first step reads.
struct FieldPtrs {
    const float* p0;
    const float* p1;
    float* out;
};

__host__ __device__ inline FieldPtrs rotate_ptrs(float* base, int64_t fe, const float* pre0, const float* pre1, int q)
{
    FieldPtrs f;
    if (q == 0) {
        f.p0 = pre0; f.p1 = pre1; f.out = base;
    } else if (q == 1) {
        f.p0 = pre1; f.p1 = base; f.out = base + fe;
    } else {
        f.p0 = base + fe * ((q - 2) & 3);
        f.p1 = base + fe * ((q - 1) & 3);
        f.out = base + fe * (q & 3);
    }
    return f;
}

}  // namespace gpvi
