// Device-side data layout and the fused time-step of the 2-D acoustic
// FDTD misfit/gradient solver (gfx950).
//
// One "unit" is one (velocity sample, shot) pair: its own three rotating
// forward fields, three adjoint fields, checkpoints, wavefield store and
// private gradient, all in HBM.  Every launch advances a set of units by one
// local time step each; every thread of a tile owns output points and does,
// for each point, the whole per-step work of the reference solver:
//
//   interior Laplacian update   fdtd2D_modelling.c:756-780
//   4 ordered ABC sweeps        fdtd2D_modelling.c:833-900
//   source / residual injection fdtd2D_modelling.c:910-911, 949-957
//   wavefield store, records    fdtd2D_modelling.c:913-937
//   imaging condition           fdtd2D_modelling.c:959-975
//
// The copy rotation of the reference (p0<-p1<-p2) becomes a rotation of
// three buffer indices, with the reference's invariant p2 == p1 at the start
// of a step turned into "a point the stencil does not update starts from p1".
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpvi {

constexpr int kMaxLc = 7;       // solver half-width (configured Lc 3..6, plus one)
constexpr int kPadRows = 16;    // margin rows above and below every field buffer
constexpr int kRowBlock = 4;    // rows per prefetch block of a strip segment

// Per-unit modes of one local step.
enum StepMode : int { kFwd = 0, kRec = 1, kAdj = 2, kIdle = 3 };

// Device unit descriptor.  Field buffers are "pitched": ntz rows of `pitch`
// floats (pitch >= ntx, a multiple of 32 floats so rows start 128-B aligned).
struct UnitDesc {
    int ntx, ntz, pitch, nx;
    int nz, pml, pml0, lc;
    int nt, k, nr, dr;
    int r_ix0, r_iz, s_ix, s_iz;
    int nseg, ckpt, last_len, nslot;   // segments of the adjoint schedule
    int tiles_x, tiles_z, start, s_off;  // tiles_x: 64-column strips, tiles_z: row segments;
                                         // local step s = global - start + s_off
    int zseg, pad2, pad3, pad4;          // rows per strip segment (one wavefront's work item)
    int total, run_fwi, grad_init, store;  // grad_init: first imaging step writes; store: keep wavefields
    float dt, dx, dz, rdt;                 // rdx/rdz/rdt = RN(1/dx) ...: exact constant division
    float rdx, rdz;
    int fastdiv, pad1;                     // 1: the host verified the reciprocal division for dx, dz, dt
    float fd[8];
    int64_t field_elems;  // stride between field buffers: (ntz + 2*kPadRows) * pitch
    const float* wtab;    // ABC taper w[i] = 1 - i/pml0, i < pml0
    // Field pointers point at row 0 of a buffer that has kPadRows rows of
    // margin before and after, so every row a strip segment touches is
    // addressable without bounds checks (margin values never reach a grid
    // point of the solution).
    float* fwd;           // 3 field buffers
    float* adj;           // 3 field buffers
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
// Schedule: local step s -> (mode, it, q, segment).
//   s in [0, nt)          : forward, it = s
//   then, segments j = nseg-1 .. 0, each:
//     (nseg > 1) recompute it = jC .. jC+Lj-1 from checkpoint j
//     adjoint   it = jC+Lj-1 .. jC
// q is the pass-local step index used for the three-buffer rotation.
struct StepInfo {
    int mode, it, q, seg;
};

__host__ __device__ inline StepInfo decode_step(const UnitDesc& u, int s)
{
    StepInfo r{kIdle, 0, 0, 0};
    if (s < 0) return r;
    if (s < u.nt) {
        r.mode = kFwd; r.it = s; r.q = s; r.seg = (u.nseg > 1) ? s / u.ckpt : 0;
        return r;
    }
    int b = s - u.nt;
    if (u.nseg == 1) {
        if (b >= u.nt) return r;
        r.mode = kAdj; r.it = u.nt - 1 - b; r.q = b; r.seg = 0;
        return r;
    }
    int j, w, lj;
    if (b < 2 * u.last_len) {
        j = u.nseg - 1; w = b; lj = u.last_len;
    } else {
        int b2 = b - 2 * u.last_len;
        j = u.nseg - 2 - b2 / (2 * u.ckpt);
        if (j < 0) return r;
        w = b2 % (2 * u.ckpt); lj = u.ckpt;
    }
    r.seg = j;
    if (w < lj) {
        r.mode = kRec; r.it = j * u.ckpt + w; r.q = w;
    } else {
        r.mode = kAdj; r.it = j * u.ckpt + (lj - 1) - (w - lj); r.q = u.nt - 1 - r.it;
    }
    return r;
}

// Three-buffer rotation with a two-field pre-history (zero, or a checkpoint):
//   q = 0: p0 = pre0, p1 = pre1, out = B0
//   q = 1: p0 = pre1, p1 = B0,   out = B1
//   q >= 2: p0 = B[(q-2)%3], p1 = B[(q-1)%3], out = B[q%3]
struct FieldPtrs {
    const float* p0;  // nullptr == all zero
    const float* p1;
    float* out;
};

__host__ __device__ inline FieldPtrs rotate_ptrs(float* base, int64_t fe, const float* pre0,
                                                 const float* pre1, int q)
{
    FieldPtrs f;
    if (q == 0) {
        f.p0 = pre0; f.p1 = pre1; f.out = base;
    } else if (q == 1) {
        f.p0 = pre1; f.p1 = base; f.out = base + fe;
    } else {
        f.p0 = base + fe * ((q - 2) % 3);
        f.p1 = base + fe * ((q - 1) % 3);
        f.out = base + fe * (q % 3);
    }
    return f;
}

}  // namespace gpvi
