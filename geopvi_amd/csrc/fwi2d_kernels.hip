// HIP kernels of the FWI-2D solver for gfx950: the fused step (single-unit
// and multi-unit persistent forms), setup and reduction kernels.
#include "fwi2d_kernels.h"
#include "fwi2d_step.h"

namespace gpvi {

// ---------------------------------------------------------------------------
// Step kernels

// One unit (the plain per-shot solver): one wavefront per strip segment,
// one time step per launch.
template <int LC, int NTH>
__global__ __launch_bounds__(NTH) void step_single_kernel(UnitDesc u, int s)
{
    constexpr int NW = NTH / 64;
    __shared__ float rows[NW][wave_lds_floats(LC, 1)];
    // readfirstlane: the compiler then knows the wave index is uniform, so
    // everything derived from it lives in scalar registers.
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int item = blockIdx.x * NW + wave;
    if (item < 2 * u.band_tz + u.tiles_x * u.tiles_z) process_item<LC, 1>(u, s, item, rows[wave]);
}

// Many units per launch, TB time steps per launch.  A grid of independent
// wavefronts (sized to what is resident at once, so no workgroup waits for
// another to finish) walks the work-item list (strip segments of every
// unit); the list is split into 8 contiguous chunks, one per XCD
// (workgroups are dispatched to XCDs round-robin), so the rows a segment
// shares with its neighbours tend to be read through the same L2.  Step and
// item count come from device memory so a captured graph segment can be
// replayed without re-instantiation.
template <int LC, int NTH, int TB>
__global__ __launch_bounds__(NTH) __attribute__((amdgpu_waves_per_eu(TB == 1 && LC <= 5 ? 6 : 1, 8))) void step_multi_kernel(const UnitDesc* __restrict__ units,
                                                         const TileRef* __restrict__ tiles,
                                                         const StepCtl* __restrict__ ctl, int step_off)
{
    constexpr int NW = NTH / 64;
    __shared__ float rows[NW][wave_lds_floats(LC, TB)];
    // readfirstlane: the wave index, the work item and the unit descriptor
    // address are then known to be uniform (scalar loads, scalar registers).
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lcs = LC - 4;
    const int ntiles = ctl->ntiles_lc[lcs];
    const int toff = ctl->tile_off[lcs];
    const int gstep = ctl->step + step_off;
    const int b = blockIdx.x;
    const int xcd = b & 7;
    const int lw = (b >> 3) * NW + wave;          // wave index within its XCD
    const int gper = (gridDim.x >> 3) * NW;       // waves per XCD
    const int chunk = (ntiles + 7) >> 3;
    const int t_begin = xcd * chunk;
    const int t_end = min(ntiles, t_begin + chunk);
    for (int t = t_begin + lw; t < t_end; t += gper) {
        const TileRef tr = tiles[toff + t];
        const UnitDesc& u = units[tr.unit];
        const int s = gstep - u.start + u.s_off;
        if (s < u.s_off || s >= u.s_off + u.total) continue;
        process_item<LC, TB>(u, s, tr.tile, rows[wave]);
    }
}

__global__ void advance_kernel(StepCtl* ctl, int n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) ctl->step = ctl->step + n;
}

__global__ void ctl_set_kernel(StepCtl* ctl, int step, int n4, int n5, int n6, int n7, int o4, int o5,
                               int o6, int o7)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->step = step;
        ctl->ntiles_lc[0] = n4; ctl->ntiles_lc[1] = n5; ctl->ntiles_lc[2] = n6; ctl->ntiles_lc[3] = n7;
        ctl->tile_off[0] = o4; ctl->tile_off[1] = o5; ctl->tile_off[2] = o6; ctl->tile_off[3] = o7;
    }
}

// ---------------------------------------------------------------------------
// Setup kernels

// Edge-replicated padding of nsample inner models (fdtd2D_modelling.c:1187-1238):
// padded(iz, ix) = inner(clamp(iz - pml), clamp(ix - pml)).
__global__ void pad_velocity_kernel(const float* __restrict__ vin, float* __restrict__ vpad, int nx,
                                    int nz, int pml, int ntx, int ntz, int pitch, int64_t fe, int nsample)
{
    const int64_t per = (int64_t)ntz * pitch;
    const int64_t total = per * nsample;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int m = (int)(i / per);
        const int64_t r = i - (int64_t)m * per;
        const int iz = (int)(r / pitch), ix = (int)(r - (int64_t)iz * pitch);
        float val = 0.f;
        if (ix < ntx) {
            const int jz = min(max(iz - pml, 0), nz - 1);
            const int jx = min(max(ix - pml, 0), nx - 1);
            val = vin[(int64_t)m * nx * nz + (int64_t)jz * nx + jx];
        }
        vpad[(int64_t)m * fe + r] = val;
    }
}

__global__ void fill_zero_kernel(float* __restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

// Per-job tables.  Ricker wavelet (fdtd2D_modelling.c:1084-1093): the
// squared argument is rounded to float, the wavelet is formed in double and
// rounded to float.  ABC taper w[i] = 1 - i/pml0 in double, stored as float
// (fdtd2D_modelling.c:227-230).
__global__ void tables_kernel(float* __restrict__ wav, int nt, float dt, float f0, float* __restrict__ wtab,
                              int pml0)
{
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it < nt) {
        const float arg = (float)it * dt - 1.f / f0;
        const double x = 3.1415926 * (double)f0 * (double)arg;
        const float tau = (float)(x * x);
        wav[it] = (float)((1.0 - 2.0 * (double)tau) * exp(-(double)tau));
    }
    if (it < pml0) wtab[it] = (float)(1.0 - 1.0 * (double)it / (double)pml0);
}

// Fill the unit table and tile list of one job from its descriptor.
__global__ void build_units_kernel(JobDesc j, UnitDesc* __restrict__ units, TileRef* __restrict__ tiles)
{
    const int nunit = j.nsample * j.ns;
    const int tpu = 2 * j.band_tz + j.tiles_x * j.tiles_z;
    const int64_t ntile = (int64_t)nunit * tpu;
    // Item order: sample, shot group, item position, shot.  The shots of one
    // sample read the same velocity tile at the same item position; next to
    // each other in the list, they run at about the same time on the same
    // XCD, so the tile comes from HBM once and from L2 for the other shots.
    // With shot_split > 0 the shots [0, shot_split) and [shot_split, ns) of a
    // sample form two groups, each with a contiguous run of the list (a batch
    // of one sample advances them as two lanes on two streams).
    const int64_t per_sample = (int64_t)j.ns * tpu;
    const int64_t split_tiles = (int64_t)j.shot_split * tpu;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ntile;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int m = (int)(i / per_sample);
        int64_t r = i - (int64_t)m * per_sample;   // position in the sample's run
        int g0 = 0, gn = j.ns;                     // the shot group: shots [g0, g0 + gn)
        if (j.shot_split > 0) {
            if (r < split_tiles) {
                gn = j.shot_split;
            } else {
                r -= split_tiles;
                g0 = j.shot_split;
                gn = j.ns - j.shot_split;
            }
        }
        const int t = (int)(r / gn);
        const int sh = g0 + (int)(r - (int64_t)t * gn);
        tiles[j.tile0 + i] = TileRef{j.unit0 + m * j.ns + sh, t};
    }
    const int un = blockIdx.x * blockDim.x + threadIdx.x;
    if (un >= nunit) return;
    const int m = un / j.ns, sh = un - (un / j.ns) * j.ns;
    UnitDesc u;
    u.ntx = j.ntx; u.ntz = j.ntz; u.pitch = j.pitch; u.nx = j.nx;
    u.nz = j.nz; u.pml = j.pml; u.pml0 = j.pml0; u.lc = j.lc;
    u.nt = j.nt; u.k = j.k; u.nr = j.nr; u.dr = j.dr;
    u.r_ix0 = j.r_ix0; u.r_iz = j.r_iz; u.s_ix = j.s_ix0 + sh * j.ds; u.s_iz = j.s_iz;
    u.nseg = j.nseg; u.ckpt = j.ckpt; u.last_len = j.last_len; u.nslot = j.nslot;
    u.tiles_x = j.tiles_x; u.tiles_z = j.tiles_z; u.start = j.start; u.s_off = j.s_off;
    u.zseg = j.zseg; u.tb = j.tb; u.stripw = j.stripw; u.band_w = j.band_w;
    u.total = j.total; u.run_fwi = j.run_fwi; u.grad_init = 1; u.store = j.store;
    u.dt = j.dt; u.dx = j.dx; u.dz = j.dz; u.rdt = j.rdt;
    u.rdx = j.rdx; u.rdz = j.rdz; u.fastdiv = j.fastdiv; u.band_tz = j.band_tz;
    u.ydx = 1.0 / (double)j.dx; u.ydz = 1.0 / (double)j.dz; u.ydt = 1.0 / (double)j.dt;
    for (int c = 0; c < 8; ++c) u.fd[c] = j.fd[c];
    u.field_elems = j.fe;
    u.wtab = j.wtab;
    u.fwd = j.fwd + (int64_t)un * j.fwd_stride;
    u.adj = j.adj + (int64_t)un * j.adj_stride;
    u.ck = j.ck ? j.ck + (int64_t)un * j.ck_stride : nullptr;
    u.save = j.save + (int64_t)un * j.save_stride;
    u.grad = j.gpriv + (int64_t)un * j.grad_stride;
    u.vel = j.vel + (int64_t)m * j.fe;
    u.zero = j.zero;
    u.wav = j.wav;
    u.syn = j.syn + ((int64_t)m * j.ns + sh) * j.nt * j.nr;
    u.obs = j.obs + (int64_t)sh * j.nt * j.nr;
    units[j.unit0 + un] = u;
}

// ---------------------------------------------------------------------------
// Reductions

// out[m][p] += sum over shots of the per-unit gradients, shot order fixed.
__global__ void grad_reduce_kernel(const float* __restrict__ gpriv, int64_t grad_stride, float* __restrict__ out,
                                   int ns, int ni, int nsample)
{
    const int64_t total = (int64_t)nsample * ni;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int m = (int)(i / ni);
        const int p = (int)(i - (int64_t)m * ni);
        double acc = (double)out[i];
        for (int s = 0; s < ns; ++s) acc += (double)gpriv[((int64_t)m * ns + s) * grad_stride + p];
        out[i] = (float)acc;
    }
}

// Per-sample partial sums of (syn - obs)^2 in double; fixed order.
constexpr int kMisfitThreads = 256;
__global__ __launch_bounds__(kMisfitThreads) void misfit_partial_kernel(
    const float* __restrict__ syn, const float* __restrict__ obs, int64_t len, int nblk,
    double* __restrict__ partial)
{
    constexpr int kWaves = kMisfitThreads / 64;
    __shared__ double wave_sum[kWaves];
    const int m = blockIdx.y, b = blockIdx.x;
    const int64_t chunk = (len + nblk - 1) / nblk;
    const int64_t lo = (int64_t)b * chunk;
    const int64_t hi = min(len, lo + chunk);
    const float* sm = syn + (int64_t)m * len;
    double acc = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kMisfitThreads) {
        const double d = (double)(sm[i] - obs[i]);
        acc += d * d;
    }
    // Wavefront butterfly (lane shuffles, a fixed order: deterministic),
    // then the block's waves in order.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((threadIdx.x & 63) == 0) wave_sum[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) t += wave_sum[w];
        partial[(int64_t)m * nblk + b] = t;
    }
}

__global__ void misfit_final_kernel(const double* __restrict__ partial, int nblk, int nsample,
                                    double* __restrict__ misfit)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nsample) return;
    double acc = 0.0;
    for (int b = 0; b < nblk; ++b) acc += partial[(int64_t)m * nblk + b];
    misfit[m] = 0.5 * acc;
}

// ---------------------------------------------------------------------------
// Host launchers

// Rows per strip segment of each launch variant.
static const int kZseg[kNumVariants] = {32, 16, 64, 8, 24, 48};

void variant_tile(int variant, int* tx, int* tz)
{
    const bool ok = variant >= 0 && variant < kNumVariants;
    *tx = ok ? kStripWidth : 0;
    *tz = ok ? kZseg[variant] : 0;
}

// `grid` = workgroups of kSingleThreads threads (one wavefront each).
hipError_t launch_step_single(int variant, int lc, int grid, hipStream_t st, const UnitDesc& u, int s)
{
    (void)variant;
    switch (lc) {
    case 4: hipLaunchKernelGGL((step_single_kernel<4, kSingleThreads>), dim3(grid), dim3(kSingleThreads), 0, st, u, s); break;
    case 5: hipLaunchKernelGGL((step_single_kernel<5, kSingleThreads>), dim3(grid), dim3(kSingleThreads), 0, st, u, s); break;
    case 6: hipLaunchKernelGGL((step_single_kernel<6, kSingleThreads>), dim3(grid), dim3(kSingleThreads), 0, st, u, s); break;
    case 7: hipLaunchKernelGGL((step_single_kernel<7, kSingleThreads>), dim3(grid), dim3(kSingleThreads), 0, st, u, s); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int LC, int TB>
static hipError_t launch_multi_lc(int grid, hipStream_t st, const UnitDesc* units, const TileRef* tiles,
                                  const StepCtl* ctl, int step_off)
{
    hipLaunchKernelGGL((step_multi_kernel<LC, kMultiThreads, TB>), dim3(grid), dim3(kMultiThreads), 0, st,
                       units, tiles, ctl, step_off);
    return hipGetLastError();
}

// `grid` = workgroups of kMultiThreads threads, a multiple of 8.
hipError_t launch_step_multi(int lc, int tb, int grid, hipStream_t st, const UnitDesc* units, const TileRef* tiles,
                             const StepCtl* ctl, int step_off)
{
    switch (lc * 2 + (tb == 2)) {
    case 8: return launch_multi_lc<4, 1>(grid, st, units, tiles, ctl, step_off);
    case 9: return launch_multi_lc<4, 2>(grid, st, units, tiles, ctl, step_off);
    case 10: return launch_multi_lc<5, 1>(grid, st, units, tiles, ctl, step_off);
    case 11: return launch_multi_lc<5, 2>(grid, st, units, tiles, ctl, step_off);
    case 12: return launch_multi_lc<6, 1>(grid, st, units, tiles, ctl, step_off);
    case 13: return launch_multi_lc<6, 2>(grid, st, units, tiles, ctl, step_off);
    case 14: return launch_multi_lc<7, 1>(grid, st, units, tiles, ctl, step_off);
    case 15: return launch_multi_lc<7, 2>(grid, st, units, tiles, ctl, step_off);
    default: return hipErrorInvalidValue;
    }
}

template <int LC, int TB>
static int resident_blocks_lc()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, step_multi_kernel<LC, kMultiThreads, TB>, kMultiThreads,
                                                     0) != hipSuccess)
        return 0;
    return nb;
}

// Workgroups of the multi-unit step resident on the current device at once
// (per-CU occupancy from the kernel's registers and LDS, times the CUs).
int multi_step_capacity(int lc, int tb)
{
    int nb = 0;
    switch (lc * 2 + (tb == 2)) {
    case 8: nb = resident_blocks_lc<4, 1>(); break;
    case 9: nb = resident_blocks_lc<4, 2>(); break;
    case 10: nb = resident_blocks_lc<5, 1>(); break;
    case 11: nb = resident_blocks_lc<5, 2>(); break;
    case 12: nb = resident_blocks_lc<6, 1>(); break;
    case 13: nb = resident_blocks_lc<6, 2>(); break;
    case 14: nb = resident_blocks_lc<7, 1>(); break;
    case 15: nb = resident_blocks_lc<7, 2>(); break;
    default: return 0;
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return nb * cus;
}

hipError_t launch_advance(hipStream_t st, StepCtl* ctl, int n)
{
    hipLaunchKernelGGL(advance_kernel, dim3(1), dim3(64), 0, st, ctl, n);
    return hipGetLastError();
}

hipError_t launch_ctl_set(hipStream_t st, StepCtl* ctl, int step, const int* ntiles4, const int* off4)
{
    hipLaunchKernelGGL(ctl_set_kernel, dim3(1), dim3(64), 0, st, ctl, step, ntiles4[0], ntiles4[1],
                       ntiles4[2], ntiles4[3], off4[0], off4[1], off4[2], off4[3]);
    return hipGetLastError();
}

hipError_t launch_zero(hipStream_t st, float* p, int64_t n)
{
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384));
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_pad_velocity(hipStream_t st, const float* vin, float* vpad, int nx, int nz, int pml,
                               int ntx, int ntz, int pitch, int64_t fe, int nsample)
{
    const int64_t total = (int64_t)ntz * pitch * nsample;
    const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(pad_velocity_kernel, dim3(grid), dim3(256), 0, st, vin, vpad, nx, nz, pml, ntx, ntz,
                       pitch, fe, nsample);
    return hipGetLastError();
}

hipError_t launch_tables(hipStream_t st, float* wav, int nt, float dt, float f0, float* wtab, int pml0)
{
    const int n = std::max(nt, pml0);
    hipLaunchKernelGGL(tables_kernel, dim3((n + 255) / 256), dim3(256), 0, st, wav, nt, dt, f0, wtab, pml0);
    return hipGetLastError();
}

hipError_t launch_build_units(hipStream_t st, const JobDesc& j, UnitDesc* units, TileRef* tiles)
{
    const int64_t ntile = (int64_t)j.nsample * j.ns * j.tiles_x * j.tiles_z;
    const int64_t nunit = (int64_t)j.nsample * j.ns;
    const int64_t work = std::max(ntile, nunit);
    const int grid = (int)std::min<int64_t>((work + 255) / 256, 65536);
    hipLaunchKernelGGL(build_units_kernel, dim3(std::max(grid, (int)((nunit + 255) / 256))), dim3(256), 0, st, j,
                       units, tiles);
    return hipGetLastError();
}

hipError_t launch_grad_reduce(hipStream_t st, const float* gpriv, int64_t grad_stride, float* out, int ns,
                              int ni, int nsample)
{
    const int64_t total = (int64_t)nsample * ni;
    const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(grad_reduce_kernel, dim3(grid), dim3(256), 0, st, gpriv, grad_stride, out, ns, ni, nsample);
    return hipGetLastError();
}

int misfit_blocks(int64_t len)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(256, len / 4096));
}

hipError_t launch_misfit(hipStream_t st, const float* syn, const float* obs, int64_t len, int nsample,
                         double* partial, double* misfit)
{
    const int nblk = misfit_blocks(len);
    hipLaunchKernelGGL(misfit_partial_kernel, dim3(nblk, nsample), dim3(kMisfitThreads), 0, st, syn, obs, len,
                       nblk, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(misfit_final_kernel, dim3((nsample + 63) / 64), dim3(64), 0, st, partial, nblk, nsample,
                       misfit);
    return hipGetLastError();
}

}  // namespace gpvi
