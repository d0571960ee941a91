// C ABI of the FWI-2D solver (declared in include/geopvi_fwi2d.h).
#include "geopvi_fwi2d.h"

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>
#include <rocprofiler-sdk-roctx/roctx.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "fwi2d_kernels.h"

using namespace gpvi;

namespace {

thread_local std::string g_err;
thread_local int g_compat_status = 0;

int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// roctx ranges (host side, visible with rocprofv3 --marker-trace): one per
// batched call, graph segment launch and engine advance; a mark per finished
// engine job.  Without a profiler attached they cost a function call.
struct TraceRange {
    explicit TraceRange(const char* fmt, ...)
    {
        char buf[160];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        roctxRangePushA(buf);
    }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess) return fail(GPVI_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

constexpr int kMaxLanes = 4;

constexpr int kDefaultVariant = 4;  // 24-row strip segments: the fastest on full batches

struct Tuning {
    int variant = -1;      // -1: per call (auto_variant); else a fixed tile variant
    int grid = 0;          // 0: automatic
    int graph_steps = 32;  // steps per graph segment
    int tb = 1;            // time steps per multi-unit launch (1 or 2)
    int lanes = 2;         // batched calls: unit groups advanced on as many streams (1..kMaxLanes)
};
Tuning g_tuning;
std::mutex g_tuning_mu;

Tuning tuning()
{
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    return g_tuning;
}

// Central second-derivative weights for solver half-width lc (configured
// Lc + 1), the table of fdtd2D_modelling.c:1272-1326 as exact rationals.
int fd_coeffs(int lc, float* fd)
{
    for (int i = 0; i < 8; ++i) fd[i] = 0.f;
    if (lc < 4 || lc > 7) return -1;
    for (int m = 0; m < lc; ++m) fd[m] = fd_weight(lc, m);
    return 0;
}

int validate(const gpvi_fwi2d_params* p)
{
    if (!p) return fail(GPVI_ERR_INVALID, "params is NULL");
    if (p->laplace_slover != 0 && p->laplace_slover != 1)
        return fail(GPVI_ERR_UNSUPPORTED, "laplace_slover=%d: 0 (finite differences) or 1 (pseudo-spectral)",
                    p->laplace_slover);
    const bool psm = p->laplace_slover == 1;  // Lc is not used by the pseudo-spectral Laplacian
    if (!psm && (p->Lc < 3 || p->Lc > 6)) return fail(GPVI_ERR_UNSUPPORTED, "Lc=%d: supported 3..6", p->Lc);
    if (p->nx < 1 || p->nz < 1 || p->ns < 1 || p->nt < 1 || p->nr < 1 || p->nt_interval < 1 || p->pml0 < 0)
        return fail(GPVI_ERR_INVALID, "nx, nz, ns, nt, nr, nt_interval must be >= 1 and pml0 >= 0");
    if (!(p->dx > 0) || !(p->dz > 0) || !(p->dt > 0) || !std::isfinite(p->dx) || !std::isfinite(p->dz) ||
        !std::isfinite(p->dt) || !std::isfinite(p->f0) || p->f0 == 0.f)
        return fail(GPVI_ERR_INVALID, "dx, dz, dt must be finite and > 0, f0 finite and non-zero");
    if (p->nr > 1 && p->dr < 1) return fail(GPVI_ERR_INVALID, "dr=%d: receivers would coincide", p->dr);
    const int lc = psm ? 0 : p->Lc + 1, pml = p->pml0 + lc;
    const long ntx = (long)p->nx + 2 * pml, ntz = (long)p->nz + 2 * pml;
    if (ntx > (1 << 20) || ntz > (1 << 20)) return fail(GPVI_ERR_INVALID, "grid too large");
    // The kernels address a field buffer (padded rows plus margins) and a
    // shot's records with 32-bit offsets.
    const long pitch = (ntx + 31 + kStripWidth - 1) / kStripWidth * kStripWidth;  // with the row shift
    // (Offsets within a field plus one field's size, and within an inner
    // grid plus one inner grid's size, must stay below 2^31.)
    if (pitch * (ntz + 2 * kPadRows) * 4 >= (1L << 30) || (long)p->nx * p->nz * 4 >= (1L << 30))
        return fail(GPVI_ERR_INVALID, "grid too large: a %ldx%ld padded field must stay under 1 GiB", ntz, ntx);
    if ((long)p->nt * p->nr >= (1L << 31))
        return fail(GPVI_ERR_INVALID, "nt*nr=%ld: a shot's records must stay under 2^31 values", (long)p->nt * p->nr);
    const long s0 = (long)pml + p->ns0, s1 = s0 + (long)(p->ns - 1) * p->ds;
    const long r0 = (long)pml + p->nr0, r1 = r0 + (long)(p->nr - 1) * p->dr;
    const long siz = (long)pml + p->depths, riz = (long)pml + p->depthr;
    if (s0 < 0 || s1 < 0 || s0 >= ntx || s1 >= ntx || siz < 0 || siz >= ntz)
        return fail(GPVI_ERR_GEOMETRY, "a source lies outside the %ldx%ld padded grid", ntz, ntx);
    if (r0 < 0 || r1 < 0 || r0 >= ntx || r1 >= ntx || riz < 0 || riz >= ntz)
        return fail(GPVI_ERR_GEOMETRY, "a receiver lies outside the %ldx%ld padded grid", ntz, ntx);
    return 0;
}

// ---------------------------------------------------------------------------
// Constant-divisor division (see cdiv_fast in fwi2d_step.h): for a divisor
// b in [2^-16, 2^16] with y = RN(1/b), q = RN(q0 + RN(a - b*q0)*y), q0 = RN(a*y),
// equals RN(a/b) for a == 0 or 2^-80 <= |a| < 2^80 (Markstein).  The kernels
// also divide that quotient by b again ((a/b)/b), whose dividend lies in
// [2^-96, 2^96].  Over a dividend range of [2^-96, 2^96] no intermediate
// under- or overflows: q0 lies in [2^-112, 2^112] (normal), and the exact
// remainder a - b*q0 is a multiple of 2^(e(q0)+e(b)-46) >= 2^-143, so the
// FMA forms it exactly even when it is subnormal.  Every rounding decision
// is then the one of the dividend scaled by a power of two into [1, 2): one
// binade, checked in full, settles the identity for the whole range; the
// binades at both ends of the range are sampled as well.  The result is
// cached per divisor.

__attribute__((target("fma"))) static bool markstein_binade_fma(float b, float y, int e, uint32_t stride)
{
    for (uint32_t m = 0; m < (1u << 23); m += stride) {
        float a;
        const uint32_t bits = (uint32_t)(127 + e) << 23 | m;
        std::memcpy(&a, &bits, 4);
        volatile float vb = b;
        const float q = a / vb;
        const float q0 = a * y;
        const float r = __builtin_fmaf(-q0, b, a);
        const float q1 = __builtin_fmaf(r, y, q0);
        if (q != q1) return false;
    }
    return true;
}

static bool markstein_binade_soft(float b, float y, int e, uint32_t stride)
{
    for (uint32_t m = 0; m < (1u << 23); m += stride) {
        float a;
        const uint32_t bits = (uint32_t)(127 + e) << 23 | m;
        std::memcpy(&a, &bits, 4);
        volatile float vb = b;
        const float q = a / vb;
        const float q0 = a * y;
        const float r = std::fma(-q0, b, a);
        const float q1 = std::fma(r, y, q0);
        if (q != q1) return false;
    }
    return true;
}

static bool markstein_ok(float b, float y)
{
    const bool hw = __builtin_cpu_supports("fma");
    auto check = [&](int e, uint32_t stride) {
        return hw ? markstein_binade_fma(b, y, e, stride) : markstein_binade_soft(b, y, e, stride * 97);
    };
    if (!check(0, 1)) return false;                           // [1, 2): every significand
    for (int e : {-96, -80, 79, 95})                          // the ends of both divisions' ranges
        if (!check(e, 61)) return false;
    return true;
}

std::mutex g_div_mu;
std::map<uint32_t, bool> g_div_ok;

bool fastdiv_ok(float b, float* recip)
{
    volatile float one = 1.f;
    *recip = one / b;
    const float ab = std::fabs(b);
    if (!(ab >= 0x1p-16f && ab <= 0x1p16f)) return false;
    uint32_t key;
    std::memcpy(&key, &b, 4);
    std::lock_guard<std::mutex> lk(g_div_mu);
    auto it = g_div_ok.find(key);
    if (it != g_div_ok.end()) return it->second;
    const bool ok = markstein_ok(b, *recip);
    g_div_ok[key] = ok;
    return ok;
}

// Field buffers: ntz rows of `pitch` floats with kPadRows margin rows before
// and after; every row starts xoff floats into its pitch (xoff < 32, so the
// first row-strip column is 128-byte aligned) and pitch = ntx + xoff rounded
// up to the 64-column strip width.  `fe` is the buffer stride and `origin`
// the offset of row 0, column 0 inside a buffer.
struct Geom {
    int lc, pml, ntx, ntz, pitch, nx, nz, ni, nt, k, nr, ns, nsave, pml0;
    int variant, tile_x, tile_z, tiles_x, tiles_z, tpu, zseg, tb, stripw, band_w, band_tz;
    int64_t fe, origin;
    float fd[8];
    float rdx, rdz, rdt;
    int fastdiv;
};

// tb: time steps per launch.  With two, a strip's 64 lanes cover its
// 64 - 2*(lc-1) owned columns plus the first step's halo on each side.
void make_geom(const gpvi_fwi2d_params* p, int variant, int tb, Geom* g)
{
    g->lc = p->laplace_slover == 1 ? 0 : p->Lc + 1;  // the pseudo-spectral path has Lc = 0 (fdtd.c:75-81)
    g->pml = p->pml0 + g->lc;
    g->ntx = p->nx + 2 * g->pml;
    g->ntz = p->nz + 2 * g->pml;
    g->nx = p->nx; g->nz = p->nz; g->ni = p->nx * p->nz;
    g->nt = p->nt; g->k = p->nt_interval; g->nr = p->nr; g->ns = p->ns;
    g->nsave = (p->nt + p->nt_interval - 1) / p->nt_interval;
    g->variant = variant;
    variant_tile(variant, &g->tile_x, &g->tile_z);
    g->zseg = g->tile_z;
    g->tb = (tb == 2 && g->lc > 0) ? 2 : 1;
    g->stripw = (g->tb == 2) ? kStripWidth - 2 * (g->lc - 1) : kStripWidth;
    g->tile_x = g->stripw;
    g->tiles_x = (g->ntx + g->tile_x - 1) / g->tile_x;
    g->tiles_z = (g->ntz + g->tile_z - 1) / g->tile_z;
    // One step per launch: the absorbing side bands (columns < pml and
    // >= ntx - pml) go to narrow side-band items, w columns wide (16 or 32,
    // >= pml; 64/w rows per wave step), so most lanes of a side-band wave
    // are band points, and the row strips over [w_left, ntx - w_right) are
    // all away from the side bands.  The two widths are chosen so the row
    // strips tile their range with the fewest idle lanes; band_w packs them
    // as w_left | w_right << 8.
    g->band_w = 0;
    g->band_tz = 0;
    if (g->tb == 1 && g->lc > 0 && g->pml <= 32) {
        int best = -1, best_waste = 1 << 30;
        for (int wl = 16; wl <= 32; wl += 16)
            for (int wr = 16; wr <= 32; wr += 16) {
                if (wl < g->pml || wr < g->pml || g->ntx < wl + wr + kStripWidth) continue;
                const int inner = g->ntx - wl - wr;
                const int n = (inner + kStripWidth - 1) / kStripWidth;
                const int waste = n * kStripWidth - inner + (wl - 16) + (wr - 16);
                if (waste < best_waste) {
                    best_waste = waste;
                    best = wl | (wr << 8);
                }
            }
        if (best >= 0) {
            const int wl = best & 0xff, wr = best >> 8;
            g->band_w = best;
            g->band_tz = g->tiles_z;
            g->tiles_x = (g->ntx - wl - wr + kStripWidth - 1) / kStripWidth;
        }
    }
    g->tpu = 2 * g->band_tz + g->tiles_x * g->tiles_z;
    // Row strips start at column w_left: shift every row by xoff floats so
    // that column lands on a 128-byte cache line (32 floats).
    const int wl0 = g->band_w & 0xff;
    const int xoff = (32 - wl0 % 32) % 32;
    g->pitch = (g->ntx + xoff + kStripWidth - 1) / kStripWidth * kStripWidth;
    g->fe = (int64_t)g->pitch * (g->ntz + 2 * kPadRows);
    g->origin = (int64_t)g->pitch * kPadRows + xoff;
    g->pml0 = p->pml0;
    fd_coeffs(g->lc, g->fd);
    const bool okx = fastdiv_ok(p->dx, &g->rdx);
    const bool okz = fastdiv_ok(p->dz, &g->rdz);
    const bool okt = fastdiv_ok(p->dt, &g->rdt);
    g->fastdiv = (okx && okz && okt) ? 1 : 0;
    const char* off = std::getenv("GPVI_NO_FASTDIV");  // every wave takes the fallback division (tests)
    if (off && off[0] && off[0] != '0') g->fastdiv = 0;
}

constexpr size_t kAlign = 256;
// The plain solver launches one unit per step: short strip segments (8 rows)
// give it the most wavefronts.
constexpr int kPlainVariant = 3;

struct Bump {
    char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count)
    {
        off = (off + kAlign - 1) / kAlign * kAlign;
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += count * sizeof(T);
        return p;
    }
};

// Adjoint schedule: full store (nseg = 1) or checkpoint segments.
struct Sched {
    int nseg, ckpt, last_len, nslot;
};

Sched make_sched(const Geom& g, int ckpt_req)
{
    Sched s;
    if (ckpt_req <= 0 || ckpt_req >= g.nt) {
        s.nseg = 1; s.ckpt = g.nt; s.last_len = g.nt; s.nslot = g.nsave;
        return s;
    }
    int c = (ckpt_req + g.k - 1) / g.k * g.k;
    s.nseg = (g.nt + c - 1) / c;
    if (s.nseg <= 1) {
        s.nseg = 1; s.ckpt = g.nt; s.last_len = g.nt; s.nslot = g.nsave;
        return s;
    }
    s.ckpt = c;
    s.last_len = g.nt - (s.nseg - 1) * c;
    s.nslot = c / g.k;
    return s;
}

// Field-buffer pointers of the layouts below point at row 0 (origin) of
// their first buffer.
struct BatchLayout {
    float *vel, *wav, *fwd, *adj, *ck, *save, *gpriv, *zero;
    double* partial;
    double* fft;    // pseudo-spectral path: per unit, ntz rows of 2 (ntx/2 + 1) doubles (in-place r2c/c2r)
    float* ktab;    // pseudo-spectral path: |k|^2 per padded point
    double* twid;   // pseudo-spectral path: (cos, sin)(2 pi m / n), m < n, for n = ntx, then for n = ntz
    UnitDesc* units;
    TileRef* tiles;
    StepCtl* ctl;
    int64_t fwd_stride, adj_stride, ck_stride, save_stride, grad_stride;
    size_t bytes;
};

// model_only: forward modelling without a wavefield store (records only);
// no adjoint fields, checkpoints, store or private gradients are laid out.
void batch_layout(const Geom& g, const Sched& sc, int nsample, bool tables, char* base, BatchLayout* L,
                  bool model_only = false)
{
    Bump b{base};
    const int64_t U = (int64_t)nsample * g.ns;
    auto rnd = [](int64_t n) { return (n + 63) / 64 * 64; };  // keep per-unit blocks 256-B aligned
    L->fwd_stride = rnd(4 * g.fe);
    L->adj_stride = model_only ? 0 : rnd(4 * g.fe);
    L->ck_stride = (sc.nseg > 1 && !model_only) ? rnd(2 * (int64_t)sc.nseg * g.fe) : 0;
    L->save_stride = model_only ? 0 : rnd((int64_t)sc.nslot * g.ni);
    L->grad_stride = model_only ? 0 : rnd(g.ni);
    auto org = [&](float* p) { return p ? p + g.origin : nullptr; };
    L->vel = org(b.take<float>((size_t)nsample * g.fe));
    L->zero = org(b.take<float>((size_t)g.fe));
    L->wav = b.take<float>(g.nt + g.pml0 + 1);  // wavelet, then the ABC taper table
    L->fwd = org(b.take<float>((size_t)(U * L->fwd_stride)));
    L->adj = org(b.take<float>((size_t)(U * L->adj_stride)));
    L->ck = L->ck_stride ? org(b.take<float>((size_t)(U * L->ck_stride))) : nullptr;
    L->save = b.take<float>((size_t)(U * L->save_stride));
    L->gpriv = b.take<float>((size_t)(U * L->grad_stride));
    L->partial = b.take<double>((size_t)nsample * misfit_blocks((int64_t)g.ns * g.nt * g.nr));
    const int64_t ntp = (int64_t)g.ntx * g.ntz;
    L->fft = g.lc == 0 ? b.take<double>((size_t)(U * g.ntz * 2 * (g.ntx / 2 + 1))) : nullptr;
    L->ktab = g.lc == 0 ? b.take<float>((size_t)ntp) : nullptr;
    L->twid = g.lc == 0 ? b.take<double>((size_t)2 * (g.ntx + g.ntz)) : nullptr;
    if (tables) {
        L->units = b.take<UnitDesc>((size_t)U);
        // sized for the shortest segments (8 rows): the workspace layout is
        // then the same for every tile variant
        L->tiles = b.take<TileRef>((size_t)(U * (2 + g.tiles_x) * ((g.ntz + 7) / 8)));
        L->ctl = b.take<StepCtl>(kMaxLanes);   // one per lane
    } else {
        L->units = nullptr; L->tiles = nullptr; L->ctl = nullptr;
    }
    L->bytes = b.off;
}

int check_workspace(void* ws, size_t have, size_t need)
{
    if (!ws) return fail(GPVI_ERR_WORKSPACE, "workspace is NULL (%zu bytes needed)", need);
    if (reinterpret_cast<uintptr_t>(ws) % kAlign)
        return fail(GPVI_ERR_WORKSPACE, "workspace must be %zu-byte aligned", kAlign);
    if (have < need) return fail(GPVI_ERR_WORKSPACE, "workspace has %zu bytes, %zu needed", have, need);
    return 0;
}

JobDesc make_job(const gpvi_fwi2d_params* p, const Geom& g, const Sched& sc, int nsample, const BatchLayout& L,
                 const float* obs, float* syn, int run_fwi, int phase, int store)
{
    JobDesc j;
    std::memset(&j, 0, sizeof j);
    j.nsample = nsample; j.ns = g.ns; j.tiles_x = g.tiles_x; j.tiles_z = g.tiles_z;
    j.zseg = g.zseg;
    j.tile0 = 0; j.unit0 = 0; j.ntx = g.ntx; j.ntz = g.ntz;
    j.pitch = g.pitch; j.nx = g.nx; j.nz = g.nz; j.pml = g.pml;
    j.pml0 = p->pml0; j.lc = g.lc; j.nt = g.nt; j.k = g.k;
    j.nr = g.nr; j.dr = (g.nr > 1) ? p->dr : 1; j.r_ix0 = g.pml + p->nr0; j.r_iz = g.pml + p->depthr;
    j.s_ix0 = g.pml + p->ns0; j.ds = p->ds; j.s_iz = g.pml + p->depths; j.nseg = sc.nseg;
    j.ckpt = sc.ckpt; j.last_len = sc.last_len; j.nslot = sc.nslot; j.start = 0;
    // Steps are counted in macro steps (launches) of g.tb time steps.
    const int fwd = macro_steps(g.nt, sc.nseg, sc.ckpt, sc.last_len, g.tb, true, false);
    const int back = macro_steps(g.nt, sc.nseg, sc.ckpt, sc.last_len, g.tb, false, true);
    if (phase == GPVI_PHASE_FORWARD) {
        j.s_off = 0; j.total = fwd;
    } else if (phase == GPVI_PHASE_ADJOINT) {
        j.s_off = fwd; j.total = back;
    } else {
        j.s_off = 0; j.total = fwd + back;
    }
    j.tb = g.tb;
    j.stripw = g.stripw;
    j.band_w = g.band_w;
    j.band_tz = g.band_tz;
    j.run_fwi = run_fwi ? 1 : 0;
    j.store = store;
    j.dt = p->dt; j.dx = p->dx; j.dz = p->dz;
    j.rdx = g.rdx; j.rdz = g.rdz; j.rdt = g.rdt; j.fastdiv = g.fastdiv;
    for (int c = 0; c < 8; ++c) j.fd[c] = g.fd[c];
    j.fe = g.fe;
    j.wtab = L.wav + g.nt;
    j.fwd_stride = L.fwd_stride; j.adj_stride = L.adj_stride; j.ck_stride = L.ck_stride;
    j.save_stride = L.save_stride; j.grad_stride = L.grad_stride;
    j.fwd = L.fwd; j.adj = L.adj; j.ck = L.ck; j.save = L.save; j.gpriv = L.gpriv;
    j.vel = L.vel; j.wav = L.wav; j.syn = syn; j.obs = obs; j.zero = L.zero;
    return j;
}

// Per-job device set-up: padded models, wavelet and taper, the zero field.
int job_setup(hipStream_t st, const gpvi_fwi2d_params* p, const Geom& g, const float* vel_inner, float* vel,
              float* wav, float* zero, int nsample)
{
    HIP_TRY(launch_pad_velocity(st, vel_inner, vel, g.nx, g.nz, g.pml, g.ntx, g.ntz, g.pitch, g.fe, nsample));
    HIP_TRY(launch_tables(st, wav, g.nt, p->dt, p->f0, wav + g.nt, g.pml0));
    HIP_TRY(launch_zero(st, zero - g.origin, g.fe));
    return 0;
}

// Workgroups of the multi-unit step resident at once on the current device
// (a multiple of 8), cached per (device, half-width, steps per launch).
int step_capacity(int lc, int tb)
{
    static std::mutex mu;
    static std::map<std::tuple<int, int, int>, int> cache;
    int dev = 0;
    (void)hipGetDevice(&dev);
    const auto key = std::make_tuple(dev, lc, tb);
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int cap = multi_step_capacity(lc, tb) / 8 * 8;
    if (cap < 8) cap = 2048;  // occupancy query unavailable: 8 waves per SIMD on 256 CUs
    cache[key] = cap;
    return cap;
}

// Workgroups of one multi-unit launch (a multiple of 8): the tuned grid if
// one is set, else one wavefront per work item, so the hardware dispatcher
// hands the next item to whichever SIMD frees up first (items differ in cost:
// absorbing bands, receiver rows, idle units).  nitems <= 0 (a graph that
// must serve any item count): twice the resident grid, each wave walking
// items until the list is done.
// Tile variant of a batched call when none is set: the longest strip
// segments (24, 16, then 8 rows) that still give each lane's launch at least
// one resident grid of work items.  Shorter segments mean more, smaller items,
// so a small batch fills the GPU; on full batches the 24-row segments are the
// fastest.  c1, two lanes: one sample 10.5 -> 16.3 samples/s with 8 rows,
// two 18.6 -> 23.1, four 28.4 -> 29.5; eight and more keep 24 rows.
int auto_variant(const gpvi_fwi2d_params* p, int nsample, const Tuning& tn)
{
    const int units = nsample * p->ns;
    const int lanes = (tn.lanes > 1 && units >= 2) ? (nsample >= 2 ? std::min(tn.lanes, nsample) : 2) : 1;
    const int per_lane = (units + lanes - 1) / lanes;
    static const int order[3] = {4, 1, 3};   // 24, 16, 8 rows per segment
    for (int v : order) {
        Geom g;
        make_geom(p, v, tn.tb, &g);
        if (g.lc == 0) return kDefaultVariant;   // pseudo-spectral: no strip items
        if ((int64_t)per_lane * g.tpu >= 4 * (int64_t)step_capacity(g.lc, g.tb)) return v;
    }
    return 3;
}

int launch_grid(int tuned, int nitems, int lc, int tb)
{
    if (tuned > 0) return std::max(8, tuned / 8 * 8);
    if (nitems <= 0) return 2 * step_capacity(lc, tb);
    const int waves_per_wg = kMultiThreads / 64;
    int g = (nitems + waves_per_wg - 1) / waves_per_wg;
    g = (g + 7) / 8 * 8;
    return std::max(8, std::min(g, 1 << 20));
}

// ---------------------------------------------------------------------------
// Graph segments: `steps` multi-unit step launches (one per half-width group
// present) followed by the step-counter advance, built by capture on a
// private stream.  Keyed by everything the captured launches bake in.
// grid 0: each half-width group's launch uses launch_grid(0, nitems, ...)
// with nitems the group's item count in `items` (0: the resident grid).
struct GraphKey {
    const void* units;
    const void* tiles;
    const void* ctl;
    int grid, tb, lc_mask, steps, device;
    int items[4];
    bool operator<(const GraphKey& o) const
    {
        return std::tie(units, tiles, ctl, grid, tb, lc_mask, steps, device, items[0], items[1], items[2],
                        items[3]) < std::tie(o.units, o.tiles, o.ctl, o.grid, o.tb, o.lc_mask, o.steps, o.device,
                                             o.items[0], o.items[1], o.items[2], o.items[3]);
    }
};

std::mutex g_graph_mu;
std::map<GraphKey, hipGraphExec_t> g_graphs;
std::map<int, hipStream_t> g_capture_streams;
constexpr size_t kMaxGraphs = 64;

// Empty the cache (after a device-wide synchronisation) when `fresh` new
// segments would not fit; called with g_graph_mu held.
int make_room_locked(size_t fresh)
{
    if (g_graphs.size() + fresh <= kMaxGraphs) return 0;
    HIP_TRY(hipDeviceSynchronize());
    for (auto& kv : g_graphs) (void)hipGraphExecDestroy(kv.second);
    g_graphs.clear();
    return 0;
}

// Find or build the segment for `key`; called with g_graph_mu held.
int get_graph_locked(const GraphKey& key, hipGraphExec_t* out)
{
    auto it = g_graphs.find(key);
    if (it != g_graphs.end()) {
        *out = it->second;
        return 0;
    }
    int rc = make_room_locked(1);
    if (rc) return rc;
    hipStream_t cap;
    auto sit = g_capture_streams.find(key.device);
    if (sit == g_capture_streams.end()) {
        HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        g_capture_streams[key.device] = cap;
    } else {
        cap = sit->second;
    }
    HIP_TRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    hipError_t err = hipSuccess;
    for (int i = 0; i < key.steps && err == hipSuccess; ++i)
        for (int lc = 4; lc <= 7 && err == hipSuccess; ++lc)
            if (key.lc_mask & (1 << (lc - 4)))
                err = launch_step_multi(lc, key.tb, launch_grid(key.grid, key.items[lc - 4], lc, key.tb), cap,
                                        (const UnitDesc*)key.units, (const TileRef*)key.tiles,
                                        (const StepCtl*)key.ctl, i);
    if (err == hipSuccess) err = launch_advance(cap, (StepCtl*)key.ctl, key.steps);
    hipGraph_t graph = nullptr;
    hipError_t e2 = hipStreamEndCapture(cap, &graph);
    if (err != hipSuccess || e2 != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return fail(GPVI_ERR_HIP, "graph capture failed: %s", hipGetErrorString(err != hipSuccess ? err : e2));
    }
    hipGraphExec_t exec;
    hipError_t e3 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e3 != hipSuccess) return fail(GPVI_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e3));
    g_graphs[key] = exec;
    *out = exec;
    return 0;
}

// Launch the segment for `key` `reps` times on `st`.  The lookup and the
// launches share one lock: a full cache is emptied (after a device-wide
// synchronisation) by whichever thread next needs a new segment, so an
// executable graph is never launched after another thread destroyed it.
int run_graph(const GraphKey& key, hipStream_t st, int reps)
{
    TraceRange tr("fwi2d step segments: %d x %d steps", reps, key.steps);
    std::lock_guard<std::mutex> lk(g_graph_mu);
    hipGraphExec_t exec;
    int rc = get_graph_locked(key, &exec);
    if (rc) return rc;
    for (int r = 0; r < reps; ++r) HIP_TRY(hipGraphLaunch(exec, st));
    return 0;
}

// The segments of `n` lanes (one graph and one stream each), enqueued in
// turn so every lane's stream has work queued from the start.
int run_graph_lanes(const GraphKey* keys, const hipStream_t* st, int n, int reps)
{
    TraceRange tr("fwi2d step segments: %d lanes x %d x %d steps", n, reps, keys[0].steps);
    std::lock_guard<std::mutex> lk(g_graph_mu);
    // Room for every lane's segment first: emptying the cache between two
    // lookups would destroy a segment already looked up.
    size_t fresh = 0;
    for (int l = 0; l < n; ++l) fresh += g_graphs.count(keys[l]) ? 0 : 1;
    int rc = make_room_locked(fresh);
    if (rc) return rc;
    hipGraphExec_t exec[kMaxLanes];
    for (int l = 0; l < n; ++l)
        if ((rc = get_graph_locked(keys[l], &exec[l]))) return rc;
    for (int r = 0; r < reps; ++r)
        for (int l = 0; l < n; ++l) HIP_TRY(hipGraphLaunch(exec[l], st[l]));
    return 0;
}

// Companion streams of the batched solver's lanes 1.. : one per (device,
// caller stream, lane), created on first use with the default priority and
// kept until the process ends.  Past kMaxCompanions of them, new callers run
// one lane.
constexpr size_t kMaxCompanions = 96;
std::mutex g_companion_mu;
std::map<std::tuple<int, hipStream_t, int>, hipStream_t> g_companions;

int companion_stream(int device, hipStream_t st, int lane, hipStream_t* out)
{
    std::lock_guard<std::mutex> lk(g_companion_mu);
    auto key = std::make_tuple(device, st, lane);
    auto it = g_companions.find(key);
    if (it != g_companions.end()) {
        *out = it->second;
        return 0;
    }
    if (g_companions.size() >= kMaxCompanions) {
        *out = nullptr;
        return 0;
    }
    hipStream_t s2;
    HIP_TRY(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
    g_companions[key] = s2;
    *out = s2;
    return 0;
}

// `to` waits for the work enqueued on `from` so far.
int stream_wait(hipStream_t to, hipStream_t from)
{
    hipEvent_t ev;
    HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e = hipEventRecord(ev, from);
    if (e == hipSuccess) e = hipStreamWaitEvent(to, ev, 0);
    (void)hipEventDestroy(ev);   // released once the recorded work completes
    if (e != hipSuccess) return fail(GPVI_ERR_HIP, "stream fork/join: %s", hipGetErrorString(e));
    return 0;
}

int current_device()
{
    int d = 0;
    (void)hipGetDevice(&d);
    return d;
}

}  // namespace

// ===========================================================================
extern "C" {

const char* gpvi_last_error(void) { return g_err.c_str(); }

int gpvi_abi_version(void) { return GPVI_ABI_VERSION; }

int gpvi_fwi2d_read_params(const char* path, gpvi_fwi2d_params* out)
{
    if (!path || !out) return fail(GPVI_ERR_INVALID, "NULL argument");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return fail(GPVI_ERR_IO, "cannot open parameter file %s", path);
    int* ints[15] = {&out->nx, &out->nz, &out->pml0, &out->Lc, &out->laplace_slover, &out->ns, &out->nt,
                     &out->ds, &out->ns0, &out->depths, &out->depthr, &out->nr, &out->dr, &out->nr0,
                     &out->nt_interval};
    float* flts[4] = {&out->dx, &out->dz, &out->dt, &out->f0};
    char line[256];
    int rc = 0;
    // Each value follows one comment line; whitespace after it is skipped.
    for (int i = 0; i < 19 && !rc; ++i) {
        if (!std::fgets(line, sizeof line, fp)) {
            rc = fail(GPVI_ERR_IO, "%s: file ends before value %d of 19", path, i + 1);
            break;
        }
        int got = (i < 15) ? std::fscanf(fp, " %d", ints[i]) : std::fscanf(fp, " %f", flts[i - 15]);
        if (got != 1) rc = fail(GPVI_ERR_IO, "%s: value %d of 19 is not a number", path, i + 1);
        else if (std::fscanf(fp, " ") < 0 && !std::feof(fp))
            rc = fail(GPVI_ERR_IO, "%s: read error", path);
    }
    std::fclose(fp);
    return rc;
}

int gpvi_fwi2d_validate(const gpvi_fwi2d_params* p) { return validate(p); }

int gpvi_fwi2d_set_tuning(int variant, int grid, int graph_steps)
{
    if (variant < -1 || variant >= kNumVariants) return fail(GPVI_ERR_INVALID, "variant %d out of range", variant);
    if (grid < 0 || graph_steps < 1 || graph_steps > 4096) return fail(GPVI_ERR_INVALID, "bad grid/graph_steps");
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    g_tuning.variant = variant;
    g_tuning.grid = grid;
    g_tuning.graph_steps = graph_steps;
    return 0;
}

int gpvi_fwi2d_set_time_blocking(int steps_per_launch)
{
    if (steps_per_launch != 1 && steps_per_launch != 2)
        return fail(GPVI_ERR_INVALID, "steps_per_launch=%d: 1 or 2", steps_per_launch);
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    g_tuning.tb = steps_per_launch;
    return 0;
}

int gpvi_fwi2d_get_time_blocking(void) { return tuning().tb; }

int gpvi_fwi2d_set_lanes(int lanes)
{
    if (lanes < 1 || lanes > kMaxLanes) return fail(GPVI_ERR_INVALID, "lanes=%d: 1 to %d", lanes, kMaxLanes);
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    g_tuning.lanes = lanes;
    return 0;
}

int gpvi_fwi2d_get_lanes(void) { return tuning().lanes; }

int gpvi_fwi2d_get_tuning(int* variant, int* grid, int* graph_steps)
{
    Tuning t = tuning();
    if (variant) *variant = t.variant;
    if (grid) *grid = t.grid;
    if (graph_steps) *graph_steps = t.graph_steps;
    return 0;
}

int gpvi_fwi2d_tile_shape(int variant, int* tile_x, int* tile_z)
{
    if (variant < 0 || variant >= kNumVariants) return fail(GPVI_ERR_INVALID, "variant %d out of range", variant);
    variant_tile(variant, tile_x, tile_z);
    return 0;
}

// ---------------------------------------------------------------------------
// Plain solver

struct PlainLayout {
    float *vel, *zero, *wav, *fwd, *adj, *save;
    size_t bytes;
};

static void plain_layout(const Geom& g, char* base, bool store, PlainLayout* L)
{
    Bump b{base};
    auto org = [&](float* p) { return p ? p + g.origin : nullptr; };
    L->vel = org(b.take<float>(g.fe));
    L->zero = org(b.take<float>(g.fe));
    L->wav = b.take<float>(g.nt + g.pml0 + 1);  // wavelet, then the ABC taper table
    L->fwd = org(b.take<float>(4 * g.fe));
    L->adj = org(b.take<float>(4 * g.fe));
    L->save = store ? b.take<float>((size_t)g.nsave * g.ni) : nullptr;
    L->bytes = b.off;
}

size_t gpvi_fwi2d_plain_workspace_bytes(const gpvi_fwi2d_params* p)
{
    if (validate(p)) return 0;
    if (p->laplace_slover == 1) return gpvi_fwi2d_batch_workspace_bytes(p, 1, 0);  // runs as a batch of one
    Geom g;
    make_geom(p, kPlainVariant, 1, &g);
    PlainLayout L;
    plain_layout(g, nullptr, true, &L);
    return L.bytes;
}

static UnitDesc plain_unit(const gpvi_fwi2d_params* p, const Geom& g)
{
    UnitDesc u;
    std::memset(&u, 0, sizeof u);
    u.ntx = g.ntx; u.ntz = g.ntz; u.pitch = g.pitch; u.nx = g.nx;
    u.nz = g.nz; u.pml = g.pml; u.pml0 = p->pml0; u.lc = g.lc;
    u.nt = g.nt; u.k = g.k; u.nr = g.nr; u.dr = (g.nr > 1) ? p->dr : 1;
    u.r_ix0 = g.pml + p->nr0; u.r_iz = g.pml + p->depthr; u.s_iz = g.pml + p->depths;
    u.nseg = 1; u.ckpt = g.nt; u.last_len = g.nt; u.nslot = g.nsave;
    u.tiles_x = g.tiles_x; u.tiles_z = g.tiles_z; u.zseg = g.zseg;
    u.band_w = g.band_w; u.band_tz = g.band_tz;
    u.dt = p->dt; u.dx = p->dx; u.dz = p->dz;
    u.rdx = g.rdx; u.rdz = g.rdz; u.rdt = g.rdt; u.fastdiv = g.fastdiv;
    u.ydx = 1.0 / (double)p->dx; u.ydz = 1.0 / (double)p->dz; u.ydt = 1.0 / (double)p->dt;
    for (int c = 0; c < 8; ++c) u.fd[c] = g.fd[c];
    u.field_elems = g.fe;
    return u;
}

static int plain_run(const gpvi_fwi2d_params* p, const float* vel_inner, float* record_syn,
                     const float* record_obs, float* grad, int run_fwi, bool adjoint, void* workspace,
                     size_t workspace_bytes, void* stream)
{
    int rc = validate(p);
    if (rc) return rc;
    if (!vel_inner || !record_syn || (adjoint && (!record_obs || !grad)))
        return fail(GPVI_ERR_INVALID, "NULL buffer");
    if (p->laplace_slover == 1)  // pseudo-spectral: a batch of one sample, shots together
        return gpvi_fwi2d_batch(p, 1, vel_inner, adjoint ? record_obs : record_syn, record_syn,
                                adjoint ? grad : nullptr, nullptr, run_fwi, 0,
                                adjoint ? GPVI_PHASE_ALL : GPVI_PHASE_FORWARD, adjoint ? 0 : GPVI_FLAG_NO_STORE,
                                workspace, workspace_bytes, stream);
    Geom g;
    make_geom(p, kPlainVariant, 1, &g);
    PlainLayout L;
    plain_layout(g, nullptr, adjoint, &L);
    if ((rc = check_workspace(workspace, workspace_bytes, L.bytes))) return rc;
    plain_layout(g, (char*)workspace, adjoint, &L);
    hipStream_t st = (hipStream_t)stream;
    if ((rc = job_setup(st, p, g, vel_inner, L.vel, L.wav, L.zero, 1))) return rc;
    UnitDesc u = plain_unit(p, g);
    u.fwd = L.fwd; u.adj = L.adj; u.save = L.save; u.grad = grad; u.vel = L.vel; u.wav = L.wav;
    u.wtab = L.wav + g.nt;
    u.zero = L.zero;
    u.run_fwi = run_fwi ? 1 : 0;
    u.grad_init = 0;            // accumulate into the caller's gradient, shot after shot
    u.store = adjoint ? 1 : 0;   // wavefield store
    u.tb = 1;
    u.stripw = g.stripw;
    u.total = macro_steps(g.nt, 1, g.nt, g.nt, 1, true, adjoint);
    const int64_t rec = (int64_t)g.nt * g.nr;
    for (int is = 0; is < g.ns; ++is) {
        u.s_ix = is * p->ds + g.pml + p->ns0;
        u.syn = record_syn + is * rec;
        u.obs = adjoint ? record_obs + is * rec : nullptr;
        for (int s = 0; s < u.total; ++s) HIP_TRY(launch_step_single(g.variant, g.lc, g.tpu, st, u, s));
    }
    return 0;
}

int gpvi_fwi2d_fwi_device(const gpvi_fwi2d_params* p, const float* vel_inner, float* record_syn,
                          const float* record_obs, float* grad, int run_fwi, void* workspace,
                          size_t workspace_bytes, void* stream)
{
    return plain_run(p, vel_inner, record_syn, record_obs, grad, run_fwi, true, workspace, workspace_bytes,
                     stream);
}

int gpvi_fwi2d_forward_device(const gpvi_fwi2d_params* p, const float* vel_inner, float* record_syn,
                              void* workspace, size_t workspace_bytes, void* stream)
{
    return plain_run(p, vel_inner, record_syn, nullptr, nullptr, 1, false, workspace, workspace_bytes, stream);
}

// ---------------------------------------------------------------------------
// Batched solver

size_t gpvi_fwi2d_batch_workspace_bytes(const gpvi_fwi2d_params* p, int nsample, int ckpt_interval)
{
    if (validate(p) || nsample < 1) return 0;
    Geom g;
    make_geom(p, kDefaultVariant, tuning().tb, &g);   // the layout is the same for every variant
    Sched sc = make_sched(g, ckpt_interval);
    BatchLayout L;
    batch_layout(g, sc, nsample, true, nullptr, &L);
    return L.bytes;
}

size_t gpvi_fwi2d_model_workspace_bytes(const gpvi_fwi2d_params* p, int nsample)
{
    if (validate(p) || nsample < 1) return 0;
    Geom g;
    make_geom(p, kDefaultVariant, tuning().tb, &g);
    Sched sc = make_sched(g, 0);
    BatchLayout L;
    batch_layout(g, sc, nsample, true, nullptr, &L, true);
    return L.bytes;
}

// ---------------------------------------------------------------------------
// Pseudo-spectral Laplacian path: batched double-precision 2-D transforms
// (hipFFT plan pairs cached per shape and device), one launch sequence per
// step.  In place: the real field of a unit has rows of rp = 2 (ntx/2 + 1)
// doubles, its spectrum ntx/2 + 1 complex values per row.
namespace {
struct PsmPlans {
    hipfftHandle fwd, inv;  // D2Z, Z2D
};
std::mutex g_psm_mu;
std::map<std::tuple<int, int, int, int, hipStream_t>, PsmPlans> g_psm_plans;

constexpr size_t kMaxPsmPlans = 16;

// Empty the plan cache (after a device-wide synchronisation) when `fresh`
// new plan pairs would not fit; called with g_psm_mu held, before the
// lookups of one call, so that no plan pair a call already holds is destroyed.
int psm_make_room(size_t fresh)
{
    if (g_psm_plans.size() + fresh <= kMaxPsmPlans) return 0;
    HIP_TRY(hipDeviceSynchronize());
    for (auto& kv : g_psm_plans) {
        (void)hipfftDestroy(kv.second.fwd);
        (void)hipfftDestroy(kv.second.inv);
    }
    g_psm_plans.clear();
    return 0;
}

// The plan pair of a shape, batch and stream, made on first use and bound to
// that stream.  One pair per stream: executions on two streams may overlap,
// and a plan's work area must not be shared by them.  Called with g_psm_mu
// held.
int psm_plans(int ntz, int ntx, int batch, hipStream_t st, PsmPlans* out)
{
    const auto key = std::make_tuple(current_device(), ntz, ntx, batch, st);
    auto it = g_psm_plans.find(key);
    if (it != g_psm_plans.end()) {
        *out = it->second;
        return 0;
    }
    if (g_psm_plans.size() >= kMaxPsmPlans)
        return fail(GPVI_ERR_HIP, "transform plan cache full (make room before the lookups)");
    const int nc = ntx / 2 + 1, rp = 2 * nc;
    int n[2] = {ntz, ntx};
    int real_embed[2] = {ntz, rp}, spec_embed[2] = {ntz, nc};
    PsmPlans pl;
    if (hipfftPlanMany(&pl.fwd, 2, n, real_embed, 1, ntz * rp, spec_embed, 1, ntz * nc, HIPFFT_D2Z, batch) !=
        HIPFFT_SUCCESS)
        return fail(GPVI_ERR_HIP, "hipfftPlanMany(D2Z %d x %d, batch %d) failed", ntz, ntx, batch);
    if (hipfftPlanMany(&pl.inv, 2, n, spec_embed, 1, ntz * nc, real_embed, 1, ntz * rp, HIPFFT_Z2D, batch) !=
        HIPFFT_SUCCESS) {
        (void)hipfftDestroy(pl.fwd);
        return fail(GPVI_ERR_HIP, "hipfftPlanMany(Z2D %d x %d, batch %d) failed", ntz, ntx, batch);
    }
    if (hipfftSetStream(pl.fwd, st) != HIPFFT_SUCCESS || hipfftSetStream(pl.inv, st) != HIPFFT_SUCCESS) {
        (void)hipfftDestroy(pl.fwd);
        (void)hipfftDestroy(pl.inv);
        return fail(GPVI_ERR_HIP, "hipfftSetStream failed");
    }
    g_psm_plans[key] = pl;
    *out = pl;
    return 0;
}

// Layout stamps: what a GPVI_PHASE_FORWARD call laid out in a workspace,
// keyed by the workspace address, so that the GPVI_PHASE_ADJOINT call on the
// same workspace uses the same layout (or fails) even if the tuning changed in
// between.  A later forward call on the same address replaces the stamp.
struct LayoutStamp {
    gpvi_fwi2d_params p;
    int nsample, ckpt, variant, tb;
    unsigned long long seq;
};
std::mutex g_stamp_mu;
std::map<const void*, LayoutStamp> g_stamps;
unsigned long long g_stamp_seq = 0;
constexpr size_t kMaxStamps = 4096;

void layout_stamp_put(const void* ws, const gpvi_fwi2d_params* p, int nsample, int ckpt, const Tuning& tn)
{
    std::lock_guard<std::mutex> lk(g_stamp_mu);
    if (g_stamps.size() >= kMaxStamps && !g_stamps.count(ws)) {
        auto oldest = g_stamps.begin();
        for (auto it = g_stamps.begin(); it != g_stamps.end(); ++it)
            if (it->second.seq < oldest->second.seq) oldest = it;
        g_stamps.erase(oldest);
    }
    g_stamps[ws] = LayoutStamp{*p, nsample, ckpt, tn.variant, tn.tb, ++g_stamp_seq};
}

int layout_stamp_get(const void* ws, const gpvi_fwi2d_params* p, int nsample, int ckpt, Tuning* tn)
{
    std::lock_guard<std::mutex> lk(g_stamp_mu);
    auto it = g_stamps.find(ws);
    if (it == g_stamps.end())
        return fail(GPVI_ERR_INVALID, "adjoint phase: this workspace was not prepared by a GPVI_PHASE_FORWARD call "
                                      "(or its record was evicted by %zu later forward calls)", kMaxStamps);
    const LayoutStamp& s = it->second;
    if (std::memcmp(&s.p, p, sizeof *p) != 0 || s.nsample != nsample || s.ckpt != ckpt)
        return fail(GPVI_ERR_INVALID, "adjoint phase: parameters, nsample or ckpt_interval differ from the forward "
                                      "phase that prepared this workspace");
    tn->variant = s.variant;
    tn->tb = s.tb;
    return 0;
}

// The pseudo-spectral steps of a batch, lane by lane: lane l advances units
// [u0[l], u0[l] + un[l]) on stream st[l] (lanes 1.. forked from st[0] by
// the caller) over its part of the transform buffer: with the hand-written
// transforms when the grid allows them (two launches per step within a
// pass, three at its first step), else with its own hipFFT plan pair.
int psm_run(const hipStream_t* st, const int* u0, const int* un, int nlane, const Geom& g, const Sched& sc,
            const JobDesc& j, const BatchLayout& L)
{
    const int64_t ustride = (int64_t)g.ntz * 2 * (g.ntx / 2 + 1);   // doubles per unit in the buffer
    // Every unit of a batch follows one schedule.  Within a pass (a run of
    // steps of one mode: forward, a segment's recompute, an adjoint stretch)
    // each step's p1 is the previous step's p2.
    UnitDesc sd;
    std::memset(&sd, 0, sizeof sd);
    sd.nt = g.nt; sd.tb = 1; sd.nseg = sc.nseg; sd.ckpt = sc.ckpt; sd.last_len = sc.last_len;
    auto mode_at = [&](int s) { return decode_step(sd, s + j.s_off).mode; };
    const char* lib_fft = std::getenv("GPVI_PSM_HIPFFT");  // the library transforms even where not needed (tests)
    if (psm_spectral_supported(g.ntx, g.ntz) && !(lib_fft && *lib_fft && std::strcmp(lib_fft, "0") != 0)) {
        for (int s = 0; s < j.total; ++s) {
            const int mode = mode_at(s);
            const bool first = s == 0 || mode_at(s - 1) != mode;
            const bool chain = s + 1 < j.total && mode_at(s + 1) == mode;
            for (int l = 0; l < nlane; ++l)
                HIP_TRY(launch_psm_spectral(st[l], L.units + u0[l], un[l], s, L.fft + u0[l] * ustride, L.ktab, L.twid,
                                            g.ntx, g.ntz, first, chain));
        }
        return 0;
    }
    PsmPlans pl[kMaxLanes];
    {
        std::lock_guard<std::mutex> lk(g_psm_mu);
        size_t fresh = 0;
        for (int l = 0; l < nlane; ++l)
            fresh += g_psm_plans.count(std::make_tuple(current_device(), g.ntz, g.ntx, un[l], st[l])) ? 0 : 1;
        int rc = psm_make_room(fresh);
        if (rc) return rc;
        for (int l = 0; l < nlane; ++l)
            if ((rc = psm_plans(g.ntz, g.ntx, un[l], st[l], &pl[l]))) return rc;
    }
    // Within a pass the point kernel leaves each step's result in the
    // transform buffer as the next step's input; the first step of a pass
    // packs its p1 from the field buffers.
    int prev_mode = -1;
    for (int s = 0; s < j.total; ++s) {
        const int mode = mode_at(s);
        for (int l = 0; l < nlane; ++l) {
            const UnitDesc* units = L.units + u0[l];
            double* buf = L.fft + u0[l] * ustride;
            auto* spec = reinterpret_cast<hipfftDoubleComplex*>(buf);
            if (mode != prev_mode) HIP_TRY(launch_psm_pack(st[l], units, un[l], s, buf, g.ntx, g.ntz));
            if (hipfftExecD2Z(pl[l].fwd, buf, spec) != HIPFFT_SUCCESS)
                return fail(GPVI_ERR_HIP, "hipfftExecD2Z failed");
            HIP_TRY(launch_psm_scale(st[l], buf, L.ktab, g.ntx, g.ntz, un[l]));
            if (hipfftExecZ2D(pl[l].inv, spec, buf) != HIPFFT_SUCCESS)
                return fail(GPVI_ERR_HIP, "hipfftExecZ2D failed");
            HIP_TRY(launch_psm_point(st[l], units, un[l], s, buf, g.ntx, g.ntz));
        }
        prev_mode = mode;
    }
    return 0;
}
}  // namespace

int gpvi_fwi2d_batch(const gpvi_fwi2d_params* p, int nsample, const float* vel_inner, const float* record_obs,
                     float* record_syn, float* grad, double* misfit, int run_fwi, int ckpt_interval, int phase,
                     int flags, void* workspace, size_t workspace_bytes, void* stream)
{
    int rc = validate(p);
    if (rc) return rc;
    if (nsample < 1) return fail(GPVI_ERR_INVALID, "nsample=%d", nsample);
    if (phase < GPVI_PHASE_ALL || phase > GPVI_PHASE_ADJOINT) return fail(GPVI_ERR_INVALID, "phase=%d", phase);
    TraceRange tr("gpvi_fwi2d_batch %dx%d x %d samples x %d shots, %d steps, phase %d, ckpt %d", p->nz, p->nx,
                  nsample, p->ns, p->nt, phase, ckpt_interval);
    if (!record_obs || !record_syn || (phase != GPVI_PHASE_ADJOINT && !vel_inner) ||
        (phase != GPVI_PHASE_FORWARD && !grad))
        return fail(GPVI_ERR_INVALID, "NULL buffer");
    Tuning tn = tuning();
    // The adjoint phase continues the forward phase's workspace: it must use
    // the layout that phase used (tile variant and steps per launch fix the
    // strip width, band items, row shift, pitch and field stride), whatever
    // the process-wide tuning says now.
    rc = phase == GPVI_PHASE_ADJOINT ? layout_stamp_get(workspace, p, nsample, ckpt_interval, &tn)
                                     : 0;
    if (rc) return rc;
    if (tn.variant < 0) tn.variant = auto_variant(p, nsample, tn);
    Geom g;
    make_geom(p, tn.variant, tn.tb, &g);
    if (g.lc == 0 && (int64_t)nsample * g.ns > 65535)  // one grid row of launches per unit
        return fail(GPVI_ERR_INVALID, "pseudo-spectral batches hold at most 65535 (sample, shot) units");
    const bool model_only = phase == GPVI_PHASE_FORWARD && (flags & GPVI_FLAG_NO_STORE);
    Sched sc = make_sched(g, model_only ? 0 : ckpt_interval);
    BatchLayout L;
    batch_layout(g, sc, nsample, true, nullptr, &L, model_only);
    if ((rc = check_workspace(workspace, workspace_bytes, L.bytes))) return rc;
    batch_layout(g, sc, nsample, true, (char*)workspace, &L, model_only);
    hipStream_t st = (hipStream_t)stream;
    if (phase != GPVI_PHASE_ADJOINT) {
        if ((rc = job_setup(st, p, g, vel_inner, L.vel, L.wav, L.zero, nsample))) return rc;
    }
    if (phase == GPVI_PHASE_FORWARD && !model_only) layout_stamp_put(workspace, p, nsample, ckpt_interval, tn);
    const int store = (phase == GPVI_PHASE_FORWARD && (flags & GPVI_FLAG_NO_STORE)) ? 0 : 1;
    JobDesc j = make_job(p, g, sc, nsample, L, record_obs, record_syn, run_fwi, phase, store);
    // Lanes: the units split into groups with contiguous runs of the tile
    // list (groups of samples, or the two shot halves of a single sample),
    // each advanced by its own launches on its own stream.
    const int per_unit = g.tpu;
    int nlane = 1;
    int lane_t0[kMaxLanes] = {0}, lane_n[kMaxLanes] = {nsample * g.ns * per_unit};
    hipStream_t lane_st[kMaxLanes] = {st};
    if (tn.lanes > 1 && !(flags & GPVI_FLAG_ONE_STREAM) && nsample * g.ns >= 2) {
        const int want = nsample >= 2 ? std::min(tn.lanes, nsample) : 2;
        int have = 1;
        while (have < want) {
            if ((rc = companion_stream(current_device(), st, have, &lane_st[have]))) return rc;
            if (!lane_st[have]) break;
            ++have;
        }
        if (have > 1) {
            nlane = have;
            if (nsample == 1) {
                j.shot_split = g.ns / 2;
                lane_n[0] = j.shot_split * per_unit;
                lane_t0[1] = lane_n[0];
                lane_n[1] = (g.ns - j.shot_split) * per_unit;
            } else {
                for (int l = 0; l < nlane; ++l) {
                    const int m0 = l * nsample / nlane, m1 = (l + 1) * nsample / nlane;
                    lane_t0[l] = m0 * g.ns * per_unit;
                    lane_n[l] = (m1 - m0) * g.ns * per_unit;
                }
            }
        }
    }
    HIP_TRY(launch_build_units(st, j, L.units, L.tiles));
    if (g.lc == 0) {
        // lanes in units (the tile counts of the finite-difference path are
        // units x tiles per unit)
        int u0[kMaxLanes], un[kMaxLanes];
        for (int l = 0; l < nlane; ++l) {
            u0[l] = lane_t0[l] / per_unit;
            un[l] = lane_n[l] / per_unit;
        }
        HIP_TRY(launch_psm_wavenumber(st, L.ktab, g.ntx, g.ntz, j.dx, j.dz));
        HIP_TRY(launch_psm_twiddles(st, L.twid, g.ntx, g.ntz));
        for (int l = 1; l < nlane; ++l)
            if ((rc = stream_wait(lane_st[l], st))) return rc;
        rc = psm_run(lane_st, u0, un, nlane, g, sc, j, L);
        for (int l = 1; l < nlane; ++l) {
            const int rj = stream_wait(st, lane_st[l]);
            if (!rc) rc = rj;
        }
        if (rc) return rc;
        if (phase != GPVI_PHASE_FORWARD)
            HIP_TRY(launch_grad_reduce(st, L.gpriv, L.grad_stride, grad, g.ns, g.ni, nsample));
        if (phase != GPVI_PHASE_ADJOINT && misfit)
            HIP_TRY(launch_misfit(st, record_syn, record_obs, (int64_t)g.ns * g.nt * g.nr, nsample, L.partial,
                                  misfit));
        return 0;
    }
    int grid[kMaxLanes];
    for (int l = 0; l < nlane; ++l) {
        int n4[4] = {0, 0, 0, 0}, o4[4] = {0, 0, 0, 0};
        n4[g.lc - 4] = lane_n[l];
        o4[g.lc - 4] = lane_t0[l];
        HIP_TRY(launch_ctl_set(st, L.ctl + l, 0, n4, o4));
        grid[l] = launch_grid(tn.grid, lane_n[l], g.lc, g.tb);
    }
    // Fork: the other lanes start after everything enqueued on the caller's
    // stream so far (job set-up, unit table, step counters).
    for (int l = 1; l < nlane; ++l)
        if ((rc = stream_wait(lane_st[l], st))) return rc;
    if (flags & GPVI_FLAG_NO_GRAPH) {
        for (int s = 0; s < j.total && !rc; ++s)
            for (int l = 0; l < nlane && !rc; ++l) {
                const hipError_t e = launch_step_multi(g.lc, g.tb, grid[l], lane_st[l], L.units, L.tiles, L.ctl + l, s);
                if (e != hipSuccess) rc = fail(GPVI_ERR_HIP, "step launch: %s", hipGetErrorString(e));
            }
    } else {
        const int seg = std::min(tn.graph_steps, j.total);
        GraphKey keys[kMaxLanes];
        for (int l = 0; l < nlane; ++l)
            keys[l] = GraphKey{L.units, L.tiles, L.ctl + l, grid[l], g.tb, 1 << (g.lc - 4), seg, current_device(),
                               {0, 0, 0, 0}};
        rc = nlane == 1 ? run_graph(keys[0], st, (j.total + seg - 1) / seg)
                        : run_graph_lanes(keys, lane_st, nlane, (j.total + seg - 1) / seg);
    }
    // Join: the caller's stream continues after every lane (also when a
    // launch failed, so no lane work is left behind the caller's stream).
    for (int l = 1; l < nlane; ++l) {
        const int rj = stream_wait(st, lane_st[l]);
        if (!rc) rc = rj;
    }
    if (rc) return rc;
    if (phase != GPVI_PHASE_FORWARD)
        HIP_TRY(launch_grad_reduce(st, L.gpriv, L.grad_stride, grad, g.ns, g.ni, nsample));
    if (phase != GPVI_PHASE_ADJOINT && misfit)
        HIP_TRY(launch_misfit(st, record_syn, record_obs, (int64_t)g.ns * g.nt * g.nr, nsample, L.partial, misfit));
    return 0;
}

// ---------------------------------------------------------------------------
// Batched solver on fixed scratch (no allocation)

}  // extern "C"

namespace {
// The scratch of gpvi_fwi2d_batch_static: a fixed array in device memory,
// one per GPU the library runs on (the code-object loader places it; it
// stays until the process ends).  In host code the array's name is only a
// stand-in for the symbol: its device address is obtained per device with
// hipGetSymbolAddress and used everywhere instead.
constexpr size_t kStaticScratchBytes = (size_t)GPVI_STATIC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch, looked up once per device.
int static_scratch(void** out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(GPVI_ERR_HIP, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(GPVI_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

extern "C" {

int gpvi_fwi2d_batch_static(const gpvi_fwi2d_params* p, int nsample, const float* vel_inner,
                            const float* record_obs, float* record_syn, float* grad, double* misfit, int run_fwi,
                            int ckpt_interval, void* stream)
{
    int rc = validate(p);
    if (rc) return rc;
    if (p->laplace_slover == 1)  // the transform plans would allocate
        return fail(GPVI_ERR_UNSUPPORTED, "gpvi_fwi2d_batch_static runs finite-difference jobs only");
    if (nsample < 1) return fail(GPVI_ERR_INVALID, "nsample=%d", nsample);
    const size_t need = gpvi_fwi2d_batch_workspace_bytes(p, nsample, ckpt_interval);
    if (need > kStaticScratchBytes)
        return fail(GPVI_ERR_WORKSPACE,
                    "gpvi_fwi2d_batch_static: this job needs %zu bytes of scratch, the fixed scratch holds %zu "
                    "(fewer samples or a checkpoint interval reduce it; gpvi_fwi2d_batch takes any workspace)",
                    need, kStaticScratchBytes);
    void* ws = nullptr;
    if ((rc = static_scratch(&ws))) return rc;
    // One launch per step on the caller's stream: an instantiated graph, a
    // companion stream or its events would be allocations.
    return gpvi_fwi2d_batch(p, nsample, vel_inner, record_obs, record_syn, grad, misfit, run_fwi, ckpt_interval,
                            GPVI_PHASE_ALL, GPVI_FLAG_NO_GRAPH | GPVI_FLAG_ONE_STREAM, ws, kStaticScratchBytes,
                            stream);
}

// ---------------------------------------------------------------------------
// Engine

}  // extern "C"

namespace {

struct EngineJob {
    long long id;
    Geom g;
    Sched sc;
    JobDesc j;
    BatchLayout L;
    int nunit;
    long long end_step;
    float* grad;
    double* misfit;
    float* syn;
    const float* obs;
    hipEvent_t done;
    bool cancelled;
};

}  // namespace

struct gpvi_engine {
    hipStream_t stream;
    int device;
    int max_units, max_tiles, grid, variant, seg_steps, tb;
    UnitDesc* units;
    TileRef* tiles;
    StepCtl* ctl;
    long long gstep;
    long long next_id;
    bool dirty;
    std::vector<EngineJob> active;
    std::map<long long, hipEvent_t> finished;
    std::vector<unsigned char> slot_used;
    std::vector<hipEvent_t> inflight;
};

namespace {

int engine_alloc_slots(gpvi_engine* e, int n)
{
    int run = 0;
    for (int i = 0; i < e->max_units; ++i) {
        run = e->slot_used[i] ? 0 : run + 1;
        if (run == n) {
            const int first = i - n + 1;
            for (int k = first; k <= i; ++k) e->slot_used[k] = 1;
            return first;
        }
    }
    return -1;
}

void engine_free_slots(gpvi_engine* e, int first, int n)
{
    for (int k = first; k < first + n; ++k) e->slot_used[k] = 0;
}

int engine_rebuild(gpvi_engine* e)
{
    int n4[4] = {0, 0, 0, 0}, o4[4] = {0, 0, 0, 0};
    for (const auto& jb : e->active) n4[jb.g.lc - 4] += jb.nunit * jb.g.tpu;
    int off = 0;
    for (int g = 0; g < 4; ++g) {
        o4[g] = off;
        off += n4[g];
    }
    int cur[4] = {o4[0], o4[1], o4[2], o4[3]};
    for (auto& jb : e->active) {
        jb.j.tile0 = cur[jb.g.lc - 4];
        cur[jb.g.lc - 4] += jb.nunit * jb.g.tpu;
        HIP_TRY(launch_build_units(e->stream, jb.j, e->units, e->tiles));
    }
    HIP_TRY(launch_ctl_set(e->stream, e->ctl, (int)e->gstep, n4, o4));
    e->dirty = false;
    return 0;
}

int engine_finalize(gpvi_engine* e, EngineJob& jb)
{
    if (!jb.cancelled) {
        HIP_TRY(launch_grad_reduce(e->stream, jb.L.gpriv, jb.L.grad_stride, jb.grad, jb.g.ns, jb.g.ni,
                                   jb.j.nsample));
        if (jb.misfit)
            HIP_TRY(launch_misfit(e->stream, jb.syn, jb.obs, (int64_t)jb.g.ns * jb.g.nt * jb.g.nr, jb.j.nsample,
                                  jb.L.partial, jb.misfit));
    }
    HIP_TRY(hipEventRecord(jb.done, e->stream));
    return 0;
}

}  // namespace

extern "C" {

int gpvi_engine_create(void* stream, int max_units, int max_tiles, gpvi_engine** out)
{
    if (!out || max_units < 1 || max_tiles < 1) return fail(GPVI_ERR_INVALID, "bad engine arguments");
    gpvi_engine* e = new gpvi_engine();
    e->stream = (hipStream_t)stream;
    e->device = current_device();
    const Tuning t = tuning();
    e->max_units = max_units;
    e->max_tiles = max_tiles;
    e->variant = t.variant >= 0 ? t.variant : kDefaultVariant;
    e->tb = t.tb;
    e->seg_steps = t.graph_steps;
    e->grid = t.grid > 0 ? std::max(8, t.grid / 8 * 8) : 0;
    e->gstep = 0;
    e->next_id = 1;
    e->dirty = true;
    e->slot_used.assign(max_units, 0);
    hipError_t err = hipMalloc((void**)&e->units, sizeof(UnitDesc) * (size_t)max_units);
    if (err == hipSuccess) err = hipMalloc((void**)&e->tiles, sizeof(TileRef) * (size_t)max_tiles);
    if (err == hipSuccess) err = hipMalloc((void**)&e->ctl, sizeof(StepCtl));
    if (err != hipSuccess) {
        if (e->units) (void)hipFree(e->units);
        if (e->tiles) (void)hipFree(e->tiles);
        if (e->ctl) (void)hipFree(e->ctl);
        delete e;
        return fail(GPVI_ERR_HIP, "engine tables: %s", hipGetErrorString(err));
    }
    *out = e;
    return 0;
}

int gpvi_engine_sync(gpvi_engine* e)
{
    if (!e) return fail(GPVI_ERR_INVALID, "NULL engine");
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (auto ev : e->inflight) (void)hipEventDestroy(ev);
    e->inflight.clear();
    return 0;
}

int gpvi_engine_destroy(gpvi_engine* e)
{
    if (!e) return 0;
    (void)hipStreamSynchronize(e->stream);
    for (auto ev : e->inflight) (void)hipEventDestroy(ev);
    for (auto& jb : e->active) (void)hipEventDestroy(jb.done);
    for (auto& kv : e->finished) (void)hipEventDestroy(kv.second);
    {
        std::lock_guard<std::mutex> lk(g_graph_mu);
        for (auto it = g_graphs.begin(); it != g_graphs.end();) {
            if (it->first.ctl == e->ctl) {
                (void)hipGraphExecDestroy(it->second);
                it = g_graphs.erase(it);
            } else {
                ++it;
            }
        }
    }
    (void)hipFree(e->units);
    (void)hipFree(e->tiles);
    (void)hipFree(e->ctl);
    delete e;
    return 0;
}

int gpvi_engine_submit(gpvi_engine* e, const gpvi_fwi2d_params* p, int nsample, const float* vel_inner,
                       const float* record_obs, float* record_syn, float* grad, double* misfit, int run_fwi,
                       int ckpt_interval, void* workspace, size_t workspace_bytes, long long* job_id)
{
    if (!e || !job_id) return fail(GPVI_ERR_INVALID, "NULL engine or job_id");
    int rc = validate(p);
    if (rc) return rc;
    if (p->laplace_slover == 1)
        return fail(GPVI_ERR_UNSUPPORTED, "the engine runs finite-difference jobs; pseudo-spectral ones go to "
                                          "gpvi_fwi2d_batch");
    if (nsample < 1 || !vel_inner || !record_obs || !record_syn || !grad)
        return fail(GPVI_ERR_INVALID, "bad job arguments");
    // The device step counter is a 32-bit int: an idle engine restarts it
    // at 0 (the next rebuild writes it in stream order after the work
    // already queued).
    if (e->active.empty()) e->gstep = 0;
    EngineJob jb;
    make_geom(p, e->variant, e->tb, &jb.g);
    jb.sc = make_sched(jb.g, ckpt_interval);
    batch_layout(jb.g, jb.sc, nsample, false, nullptr, &jb.L);
    if ((rc = check_workspace(workspace, workspace_bytes, jb.L.bytes))) return rc;
    batch_layout(jb.g, jb.sc, nsample, false, (char*)workspace, &jb.L);
    jb.nunit = nsample * jb.g.ns;
    int tiles_now = 0;
    for (const auto& a : e->active) tiles_now += a.nunit * a.g.tpu;
    if (tiles_now + jb.nunit * jb.g.tpu > e->max_tiles)
        return fail(GPVI_ERR_CAPACITY, "tile list full (%d + %d > %d)", tiles_now, jb.nunit * jb.g.tpu, e->max_tiles);
    const int first = engine_alloc_slots(e, jb.nunit);
    if (first < 0) return fail(GPVI_ERR_CAPACITY, "no run of %d free unit slots", jb.nunit);
    hipError_t err = hipEventCreateWithFlags(&jb.done, hipEventDisableTiming);
    if (err != hipSuccess) {
        engine_free_slots(e, first, jb.nunit);
        return fail(GPVI_ERR_HIP, "hipEventCreate: %s", hipGetErrorString(err));
    }
    if ((rc = job_setup(e->stream, p, jb.g, vel_inner, jb.L.vel, jb.L.wav, jb.L.zero, nsample))) {
        engine_free_slots(e, first, jb.nunit);
        (void)hipEventDestroy(jb.done);
        return rc;
    }
    jb.j = make_job(p, jb.g, jb.sc, nsample, jb.L, record_obs, record_syn, run_fwi, GPVI_PHASE_ALL, 1);
    if (e->gstep + jb.j.total + e->seg_steps > INT_MAX) {
        engine_free_slots(e, first, jb.nunit);
        (void)hipEventDestroy(jb.done);
        return fail(GPVI_ERR_CAPACITY, "step counter at %lld: let the engine drain before submitting", e->gstep);
    }
    jb.j.unit0 = first;
    jb.j.start = (int)e->gstep;
    jb.end_step = e->gstep + jb.j.total;
    jb.grad = grad;
    jb.misfit = misfit;
    jb.syn = record_syn;
    jb.obs = record_obs;
    jb.cancelled = false;
    jb.id = e->next_id++;
    e->active.push_back(jb);
    e->dirty = true;
    *job_id = jb.id;
    return 0;
}

int gpvi_engine_advance(gpvi_engine* e, int nsteps)
{
    if (!e) return fail(GPVI_ERR_INVALID, "NULL engine");
    TraceRange tr("gpvi_engine_advance %zu jobs", e->active.size());
    int rc;
    if (e->dirty && (rc = engine_rebuild(e))) return rc;
    if (e->active.empty() || nsteps <= 0) return 0;
    int mask = 0;
    for (const auto& jb : e->active) mask |= 1 << (jb.g.lc - 4);
    // Items 0: every launch uses the resident grid, so the graph survives
    // changes of the job set.
    GraphKey key{e->units, e->tiles, e->ctl, e->grid, e->tb, mask, e->seg_steps, e->device, {0, 0, 0, 0}};
    long long pending = 0;  // never run past the last active job's last step
    for (const auto& jb : e->active) pending = std::max(pending, jb.end_step - e->gstep);
    const long long todo = std::min<long long>(nsteps, pending);
    const int reps = (int)((todo + e->seg_steps - 1) / e->seg_steps);
    for (int r = 0; r < reps; ++r) {
        // Keep at most three segments queued ahead of the device.
        while (e->inflight.size() >= 3) {
            HIP_TRY(hipEventSynchronize(e->inflight.front()));
            (void)hipEventDestroy(e->inflight.front());
            e->inflight.erase(e->inflight.begin());
        }
        if ((rc = run_graph(key, e->stream, 1))) return rc;
        hipEvent_t ev;
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        HIP_TRY(hipEventRecord(ev, e->stream));
        e->inflight.push_back(ev);
        e->gstep += e->seg_steps;
    }
    int ncomplete = 0;
    for (size_t i = 0; i < e->active.size();) {
        EngineJob& jb = e->active[i];
        if (jb.end_step <= e->gstep) {
            if ((rc = engine_finalize(e, jb))) return rc;
            {
                char msg[64];
                std::snprintf(msg, sizeof msg, "engine job %lld finalised", jb.id);
                roctxMarkA(msg);
            }
            engine_free_slots(e, jb.j.unit0, jb.nunit);
            e->finished[jb.id] = jb.done;
            e->active.erase(e->active.begin() + i);
            e->dirty = true;
            ++ncomplete;
        } else {
            ++i;
        }
    }
    return ncomplete;
}

int gpvi_engine_job_status(gpvi_engine* e, long long job_id)
{
    if (!e) return fail(GPVI_ERR_INVALID, "NULL engine");
    auto it = e->finished.find(job_id);
    if (it != e->finished.end()) {
        hipError_t q = hipEventQuery(it->second);
        if (q == hipSuccess) return 1;
        if (q == hipErrorNotReady) return 0;
        return fail(GPVI_ERR_HIP, "hipEventQuery: %s", hipGetErrorString(q));
    }
    for (const auto& jb : e->active)
        if (jb.id == job_id) return 0;
    return fail(GPVI_ERR_UNKNOWN_JOB, "unknown job %lld", job_id);
}

int gpvi_engine_release(gpvi_engine* e, long long job_id)
{
    if (!e) return fail(GPVI_ERR_INVALID, "NULL engine");
    auto it = e->finished.find(job_id);
    if (it == e->finished.end()) return fail(GPVI_ERR_UNKNOWN_JOB, "job %lld is not finished", job_id);
    HIP_TRY(hipEventSynchronize(it->second));
    (void)hipEventDestroy(it->second);
    e->finished.erase(it);
    return 0;
}

int gpvi_engine_cancel(gpvi_engine* e, long long job_id)
{
    if (!e) return fail(GPVI_ERR_INVALID, "NULL engine");
    for (size_t i = 0; i < e->active.size(); ++i) {
        EngineJob& jb = e->active[i];
        if (jb.id != job_id) continue;
        jb.cancelled = true;
        int rc = engine_finalize(e, jb);  // records the event after all work already queued
        if (rc) return rc;
        engine_free_slots(e, jb.j.unit0, jb.nunit);
        e->finished[jb.id] = jb.done;
        e->active.erase(e->active.begin() + i);
        e->dirty = true;
        return 0;
    }
    return fail(GPVI_ERR_UNKNOWN_JOB, "job %lld is not active", job_id);
}

int gpvi_engine_active_jobs(gpvi_engine* e) { return e ? (int)e->active.size() : 0; }

int gpvi_engine_active_units(gpvi_engine* e)
{
    if (!e) return 0;
    int n = 0;
    for (const auto& jb : e->active) n += jb.nunit;
    return n;
}

int gpvi_engine_pending_steps(gpvi_engine* e)
{
    if (!e) return 0;
    long long m = 0;
    for (const auto& jb : e->active) m = std::max(m, jb.end_step - e->gstep);
    return (int)m;
}

// ---------------------------------------------------------------------------
// Reference-compatible host entry points

// Checkpoint interval of a one-sample call: 0 (the full wavefield store) when
// its workspace fits in 3/4 of the free device memory, else the interval
// that minimises a unit's memory, C* = sqrt(2 nt k ntp / ni) rounded to a
// multiple of k (0 once it reaches nt).
static int compat_ckpt(const gpvi_fwi2d_params* p)
{
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) free_b = 0;
    if ((double)gpvi_fwi2d_batch_workspace_bytes(p, 1, 0) <= 0.75 * (double)free_b) return 0;
    const int lc = p->laplace_slover == 1 ? 0 : p->Lc + 1, pml = p->pml0 + lc, k = p->nt_interval;
    const double ntp = (double)(p->nx + 2 * pml) * (double)(p->nz + 2 * pml), ni = (double)p->nx * (double)p->nz;
    const double c = std::sqrt(2.0 * p->nt * k * ntp / ni);
    const int ci = std::max(k, (int)std::lround(c / k) * k);
    return ci < p->nt ? ci : 0;
}

// One sample through the batched solver (its shots advance together), host
// buffers in and out; fwi_2D adds to the caller's gradient, as the
// reference does.
static int compat_run(char input_file[200], float* vel_inner, float* record_syn, float* record_obs, float* grad,
                      int run_fwi, bool adjoint, int verbose)
{
    gpvi_fwi2d_params p;
    int rc = gpvi_fwi2d_read_params(input_file, &p);
    if (rc) return rc;
    if ((rc = validate(&p))) return rc;
    if (verbose)
        std::printf("fwi2d: nx=%d nz=%d nt=%d ns=%d dx=%f dz=%f dt=%f\n", p.nx, p.nz, p.nt, p.ns, p.dx, p.dz, p.dt);
    const size_t ni = (size_t)p.nx * p.nz, nrec = (size_t)p.ns * p.nt * p.nr;
    const int ckpt = adjoint ? compat_ckpt(&p) : 0;
    const size_t ws = adjoint ? gpvi_fwi2d_batch_workspace_bytes(&p, 1, ckpt) : gpvi_fwi2d_model_workspace_bytes(&p, 1);
    float *d_vel = nullptr, *d_syn = nullptr, *d_obs = nullptr, *d_grad = nullptr;
    double* d_mis = nullptr;
    void* d_ws = nullptr;
    hipError_t err = hipMalloc((void**)&d_vel, ni * sizeof(float));
    if (err == hipSuccess) err = hipMalloc((void**)&d_syn, nrec * sizeof(float));
    if (err == hipSuccess) err = hipMalloc((void**)&d_obs, nrec * sizeof(float));
    if (err == hipSuccess && adjoint) err = hipMalloc((void**)&d_grad, ni * sizeof(float));
    if (err == hipSuccess && adjoint) err = hipMalloc((void**)&d_mis, sizeof(double));
    if (err == hipSuccess) err = hipMalloc(&d_ws, ws);
    if (err == hipSuccess) err = hipMemcpy(d_vel, vel_inner, ni * sizeof(float), hipMemcpyHostToDevice);
    if (err == hipSuccess)
        err = adjoint ? hipMemcpy(d_obs, record_obs, nrec * sizeof(float), hipMemcpyHostToDevice)
                      : hipMemset(d_obs, 0, nrec * sizeof(float));
    if (err == hipSuccess && adjoint) err = hipMemcpy(d_grad, grad, ni * sizeof(float), hipMemcpyHostToDevice);
    if (err != hipSuccess) rc = fail(GPVI_ERR_HIP, "fwi_2D setup: %s", hipGetErrorString(err));
    if (!rc) {
        rc = adjoint ? gpvi_fwi2d_batch(&p, 1, d_vel, d_obs, d_syn, d_grad, d_mis, run_fwi, ckpt, GPVI_PHASE_ALL, 0,
                                        d_ws, ws, nullptr)
                     : gpvi_fwi2d_batch(&p, 1, d_vel, d_obs, d_syn, nullptr, nullptr, 1, 0, GPVI_PHASE_FORWARD,
                                        GPVI_FLAG_NO_STORE, d_ws, ws, nullptr);
    }
    if (!rc) {
        err = hipDeviceSynchronize();
        if (err == hipSuccess) err = hipMemcpy(record_syn, d_syn, nrec * sizeof(float), hipMemcpyDeviceToHost);
        if (err == hipSuccess && adjoint) err = hipMemcpy(grad, d_grad, ni * sizeof(float), hipMemcpyDeviceToHost);
        if (err != hipSuccess) rc = fail(GPVI_ERR_HIP, "fwi_2D: %s", hipGetErrorString(err));
    }
    (void)hipFree(d_vel); (void)hipFree(d_syn); (void)hipFree(d_obs); (void)hipFree(d_grad); (void)hipFree(d_mis);
    (void)hipFree(d_ws);
    return rc;
}

void fwi_2D(char input_file[200], float* vel_inner, float* record_syn, float* record_obs, float* grad, int run_fwi,
            int verbose)
{
    g_compat_status = compat_run(input_file, vel_inner, record_syn, record_obs, grad, run_fwi, true, verbose);
    if (g_compat_status) std::fprintf(stderr, "fwi_2D: %s\n", g_err.c_str());
}

void forward_2D(char input_file[200], float* vel_inner, float* record_syn, int run_fwi, int verbose)
{
    (void)run_fwi;
    g_compat_status = compat_run(input_file, vel_inner, record_syn, nullptr, nullptr, 1, false, verbose);
    if (g_compat_status) std::fprintf(stderr, "forward_2D: %s\n", g_err.c_str());
}

int gpvi_compat_last_status(void) { return g_compat_status; }

}  // extern "C"
