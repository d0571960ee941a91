// Host-side view of the FWI-2D kernels (launchers and job descriptors).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "fwi2d_device.h"

namespace gpvi {

// Everything needed to build the unit table of one job (nsample samples x
// ns shots) on the device; passed by value as a kernel argument so no host
// buffer has to outlive an asynchronous copy.
struct JobDesc {
    int nsample, ns, tiles_x, tiles_z;
    int tile0, unit0, ntx, ntz;
    int pitch, nx, nz, pml;
    int pml0, lc, nt, k;
    int nr, dr, r_ix0, r_iz;
    int s_ix0, ds, s_iz, nseg;
    int ckpt, last_len, nslot, start;
    int s_off, total, run_fwi, store;
    int zseg, tb, stripw, band_w;
    float dt, dx, dz, rdt;
    float rdx, rdz;
    int fastdiv, band_tz;
    int shot_split;       // > 0: a sample's shots [0, shot_split) and [shot_split, ns) list their tiles apart
    float fd[8];
    int64_t fe;
    const float* wtab;
    int64_t fwd_stride, adj_stride, ck_stride, save_stride, grad_stride;
    float* fwd;
    float* adj;
    float* ck;
    float* save;
    float* gpriv;
    const float* vel;
    const float* wav;
    float* syn;
    const float* obs;
    const float* zero;
};

// Launch variants = rows per strip segment (the work item of one wavefront);
// strips are always 64 columns wide.
void variant_tile(int variant, int* tx, int* tz);
constexpr int kNumVariants = 6;
constexpr int kStripWidth = 64;
constexpr int kMultiThreads = 256;   // workgroup of the multi-unit step (4 independent waves)
constexpr int kSingleThreads = 64;   // workgroup of the single-unit step (1 wave)

hipError_t launch_step_single(int variant, int lc, int grid, hipStream_t st, const UnitDesc& u, int s);
hipError_t launch_step_multi(int lc, int tb, int grid, hipStream_t st, const UnitDesc* units, const TileRef* tiles,
                             const StepCtl* ctl, int step_off);
int multi_step_capacity(int lc, int tb);
hipError_t launch_advance(hipStream_t st, StepCtl* ctl, int n);
hipError_t launch_ctl_set(hipStream_t st, StepCtl* ctl, int step, const int* ntiles4, const int* off4);
hipError_t launch_zero(hipStream_t st, float* p, int64_t n);
hipError_t launch_pad_velocity(hipStream_t st, const float* vin, float* vpad, int nx, int nz, int pml,
                               int ntx, int ntz, int pitch, int64_t fe, int nsample);
hipError_t launch_tables(hipStream_t st, float* wav, int nt, float dt, float f0, float* wtab, int pml0);
hipError_t launch_build_units(hipStream_t st, const JobDesc& j, UnitDesc* units, TileRef* tiles);
hipError_t launch_grad_reduce(hipStream_t st, const float* gpriv, int64_t grad_stride, float* out, int ns,
                              int ni, int nsample);
int misfit_blocks(int64_t len);

// Pseudo-spectral Laplacian path (fwi2d_psm.hip).
hipError_t launch_psm_wavenumber(hipStream_t st, float* k, int ntx, int ntz, float dx, float dz);
hipError_t launch_psm_pack(hipStream_t st, const UnitDesc* units, int nunit, int s, double* buf, int ntx, int ntz);
hipError_t launch_psm_scale(hipStream_t st, double* buf, const float* k, int ntx, int ntz, int nunit);
hipError_t launch_psm_point(hipStream_t st, const UnitDesc* units, int nunit, int s, double* buf, int ntx,
                            int ntz);
// The hand-written transforms: lengths ntx and ntz factor into 2, 3, 5, 7,
// 11, 13 and fit the LDS tiling.  One step: the row-forward launch when it
// starts a pass (first), the column launch, and the row-backward launch,
// which also does the next step's row-forward work when the next step
// continues the pass (chain).
bool psm_spectral_supported(int ntx, int ntz);
hipError_t launch_psm_twiddles(hipStream_t st, double* tw, int ntx, int ntz);
hipError_t launch_psm_spectral(hipStream_t st, const UnitDesc* units, int nunit, int s, double* buf,
                               const float* k, const double* tw, int ntx, int ntz, bool first, bool chain);
hipError_t launch_misfit(hipStream_t st, const float* syn, const float* obs, int64_t len, int nsample,
                         double* partial, double* misfit);

}  // namespace gpvi
