/*
 * geopvi_fwi2d.h -- C ABI of the MI355X (gfx950) FWI-2D misfit/gradient
 * solver, libgeopvi_fwi2d.so.
 *
 * The library replaces the native layer of GeoPVI's 2-D acoustic FWI
 * (geopvi/forward/fwi2d/fdtd2D_modelling.c, bound by pyacofwi2D.pyx).
 * Plain C types only: pointers and sizes, no torch or HIP types in the
 * signatures.  `stream` arguments are hipStream_t values passed as void*
 * (NULL = the default stream).  Device-pointer entry points enqueue work on
 * that stream and return without synchronising; host-pointer entry points
 * (fwi_2D, forward_2D) are synchronous.
 *
 * Status codes: 0 on success, a negative GPVI_ERR_* on failure; the
 * message of the last failure on the calling thread is gpvi_last_error().
 * Nothing in this library calls exit().
 */
#ifndef GEOPVI_FWI2D_H
#define GEOPVI_FWI2D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPVI_ABI_VERSION 1

#define GPVI_OK 0
#define GPVI_ERR_INVALID (-1)     /* bad argument or parameter value */
#define GPVI_ERR_UNSUPPORTED (-2) /* laplace_slover not 0/1, Lc outside 3..6, PSM job to the engine */
#define GPVI_ERR_GEOMETRY (-3)    /* a source or receiver outside the padded grid */
#define GPVI_ERR_IO (-4)          /* parameter file missing or malformed */
#define GPVI_ERR_HIP (-5)         /* HIP runtime failure */
#define GPVI_ERR_WORKSPACE (-6)   /* workspace missing, too small or misaligned */
#define GPVI_ERR_CAPACITY (-7)    /* engine unit/tile tables full */
#define GPVI_ERR_UNKNOWN_JOB (-8) /* engine job id not known */

/* Phases of a batched evaluation. */
#define GPVI_PHASE_ALL 0      /* forward + adjoint in one call */
#define GPVI_PHASE_FORWARD 1  /* forward only: records, misfit, checkpoints kept in the workspace */
#define GPVI_PHASE_ADJOINT 2  /* adjoint from a workspace left by GPVI_PHASE_FORWARD: same parameters,
                                 nsample and ckpt_interval, and the launch layout that call used (a later
                                 change of tuning or time blocking does not apply); GPVI_ERR_INVALID if the
                                 workspace was not prepared by such a call */

/* Flags of gpvi_fwi2d_batch. */
#define GPVI_FLAG_NO_GRAPH 1  /* one launch per step instead of replayed graph segments */
#define GPVI_FLAG_NO_STORE 2  /* forward modelling only: no wavefield store (PHASE_FORWARD) */
#define GPVI_FLAG_ONE_STREAM 4 /* all launches on the caller's stream (no second lane; see
                                  gpvi_fwi2d_set_lanes) */

/*
 * Run parameters: the 19 values of the reference's positional parameter
 * file, in the same order (fdtd2D_modelling.c:1337-1356; written by
 * geopvi/forward/fwi2d/posterior.py:27-68).  Lc is the configured
 * half-order 3..6; the solver uses Lc + 1 (fdtd2D_modelling.c:70).
 */
typedef struct gpvi_fwi2d_params {
    int nx, nz, pml0, Lc, laplace_slover, ns, nt, ds, ns0;
    int depths, depthr, nr, dr, nr0, nt_interval;
    float dx, dz, dt, f0;
} gpvi_fwi2d_params;

const char *gpvi_last_error(void);
int gpvi_abi_version(void);

/* Parse the positional parameter file.  Replaces read_parameters
 * (fdtd2D_modelling.c:1328-1379); returns GPVI_ERR_IO instead of exit(0). */
int gpvi_fwi2d_read_params(const char *path, gpvi_fwi2d_params *out);

/* Check that these parameters can run (no counterpart in the reference,
 * which has no checks): laplace_slover 0 (finite differences, Lc in 3..6) or
 * 1 (pseudo-spectral Laplacian, Lc unused and pml = pml0 as in
 * fdtd2D_modelling.c:75-85), sizes positive, every source and receiver
 * inside the padded grid, a padded field under 1 GiB and a shot's records
 * under 2^31 values (the kernels use 32-bit offsets; GPVI_ERR_INVALID
 * otherwise).  Every entry point below runs both Laplacians
 * except the engine, which takes finite-difference jobs only; a
 * pseudo-spectral job runs as one batched call (double-precision transforms,
 * pstd_2d_calculate_p fdtd2D_modelling.c:783-830: the library's own
 * mixed-radix kernels when ntx and ntz factor into 2, 3, 5, 7, 11, 13 and
 * are at most 1638 (1365 if odd), hipFFT otherwise or with GPVI_PSM_HIPFFT=1
 * set). */
int gpvi_fwi2d_validate(const gpvi_fwi2d_params *p);

/* ---- Plain solver: one sample, shots in sequence, one launch per step ----
 * Device-memory form of fwi_2D (fdtd2D_modelling.c:18-353): same buffers and
 * layouts -- vel_inner [nz][nx], record_syn/record_obs [ns][nt][nr],
 * grad [nz][nx] ACCUMULATED (+=, the caller zeroes it, as pyacofwi2D.pyx:25
 * does).  The workspace (device memory, 256-B aligned) holds the padded
 * model, fields and the full wavefield store. */
size_t gpvi_fwi2d_plain_workspace_bytes(const gpvi_fwi2d_params *p);
int gpvi_fwi2d_fwi_device(const gpvi_fwi2d_params *p, const float *vel_inner, float *record_syn,
                          const float *record_obs, float *grad, int run_fwi, void *workspace,
                          size_t workspace_bytes, void *stream);

/* Device-memory form of forward_2D (fdtd2D_modelling.c:356-675): records of
 * every shot; no wavefield store. */
int gpvi_fwi2d_forward_device(const gpvi_fwi2d_params *p, const float *vel_inner, float *record_syn,
                              void *workspace, size_t workspace_bytes, void *stream);

/* ---- Batched solver: nsample samples x ns shots per launch ----
 * Replaces the per-sample process fan-out around fwi_2D
 * (geopvi/forward/fwi2d/posterior.py:131-149) with one launch per time step
 * over every (sample, shot) unit, replayed as HIP graph segments.
 *   vel_inner  [nsample][nz][nx]   record_obs [ns][nt][nr] (shared)
 *   record_syn [nsample][ns][nt][nr]   grad [nsample][nz][nx] (+=)
 *   misfit     [nsample] double, 0.5*sum((syn-obs)^2) (posterior.py:109)
 * ckpt_interval <= 0 or >= nt keeps every nt_interval-th wavefield (the
 * reference's store); otherwise (p0, p1) checkpoints are kept every
 * ckpt_interval steps (rounded up to a multiple of nt_interval) and each
 * segment is recomputed during the adjoint.  misfit may be NULL; grad may be
 * NULL in GPVI_PHASE_FORWARD. */
size_t gpvi_fwi2d_batch_workspace_bytes(const gpvi_fwi2d_params *p, int nsample, int ckpt_interval);
/* Workspace of a forward-modelling-only call (GPVI_PHASE_FORWARD with
 * GPVI_FLAG_NO_STORE; the reference's forward_2D, fdtd2D_modelling.c:356-675,
 * allocates a wavefield store it never reads): fields and records only. */
size_t gpvi_fwi2d_model_workspace_bytes(const gpvi_fwi2d_params *p, int nsample);
int gpvi_fwi2d_batch(const gpvi_fwi2d_params *p, int nsample, const float *vel_inner,
                     const float *record_obs, float *record_syn, float *grad, double *misfit,
                     int run_fwi, int ckpt_interval, int phase, int flags, void *workspace,
                     size_t workspace_bytes, void *stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  gpvi_fwi2d_batch with GPVI_PHASE_ALL and one launch
 * per step, its workspace taken from a fixed device array of
 * GPVI_STATIC_SCRATCH_BYTES per GPU (held from library load until the
 * process ends) instead of an argument: the caller passes inputs and outputs
 * only.  Nothing is allocated.  A job whose workspace
 * (gpvi_fwi2d_batch_workspace_bytes) exceeds the scratch is refused with
 * GPVI_ERR_WORKSPACE; pseudo-spectral jobs with GPVI_ERR_UNSUPPORTED.  Every
 * call on a GPU uses that GPU's one scratch array: calls that overlap on one
 * GPU (two streams or two threads) race and must be serialised by the
 * caller. */
#define GPVI_STATIC_SCRATCH_BYTES (64u << 20)
int gpvi_fwi2d_batch_static(const gpvi_fwi2d_params *p, int nsample, const float *vel_inner,
                            const float *record_obs, float *record_syn, float *grad, double *misfit, int run_fwi,
                            int ckpt_interval, void *stream);

/* Launch shape of the multi-unit step: tile variant (0..5: rows per strip
 * segment; -1, the default: chosen per batched call, the longest segments
 * that still give each lane at least one resident grid of work items),
 * persistent grid (a multiple of 8; 0 = one wavefront per work item) and
 * steps per graph segment.  Process-wide; the autotuner sets them.  The
 * workspace sizes do not depend on the variant. */
int gpvi_fwi2d_set_tuning(int variant, int grid, int graph_steps);
int gpvi_fwi2d_get_tuning(int *variant, int *grid, int *graph_steps);
int gpvi_fwi2d_tile_shape(int variant, int *tile_x, int *tile_z);
/* Time steps per multi-unit launch (1 or 2; default 1).  With 2, each
 * launch advances every unit by two consecutive steps of its pass, the
 * second computed from the first in registers/LDS (temporal blocking). */
int gpvi_fwi2d_set_time_blocking(int steps_per_launch);
int gpvi_fwi2d_get_time_blocking(void);
/* Lanes of a batched call (1 to 4; default 2).  With n > 1, the (sample,
 * shot) units are split into n groups of samples (a single sample: its two
 * shot halves), each advanced by its own step launches on its own stream:
 * the caller's stream and companion streams the library creates once per
 * caller stream (default priority), joined back to the caller's stream by
 * events.  The step launches of one group then fill the GPU while those of
 * another drain.  Streams only overlap on hardware queues of their own
 * (GPU_MAX_HW_QUEUES).  Results are the same bits for every n.
 * Process-wide; one lane is always used with GPVI_FLAG_ONE_STREAM, for a
 * single unit and by gpvi_fwi2d_batch_static. */
int gpvi_fwi2d_set_lanes(int lanes);
int gpvi_fwi2d_get_lanes(void);

/* ---- Engine: continuous batching of misfit+gradient jobs ----
 * Jobs of any grid size join the running set at a step boundary; every step
 * launch covers the units of all active jobs; a job's outputs are final when
 * gpvi_engine_job_status() reports 1.  Caller-owned buffers (inputs, outputs,
 * job workspace of gpvi_fwi2d_batch_workspace_bytes bytes) must stay valid
 * until then.  One host thread drives an engine. */
typedef struct gpvi_engine gpvi_engine;
int gpvi_engine_create(void *stream, int max_units, int max_tiles, gpvi_engine **out);
int gpvi_engine_destroy(gpvi_engine *e);
int gpvi_engine_submit(gpvi_engine *e, const gpvi_fwi2d_params *p, int nsample, const float *vel_inner,
                       const float *record_obs, float *record_syn, float *grad, double *misfit,
                       int run_fwi, int ckpt_interval, void *workspace, size_t workspace_bytes,
                       long long *job_id);
/* Advance every active job by nsteps steps (rounded up to whole graph
 * segments); returns the number of jobs that completed (their finalising
 * kernels are enqueued). */
int gpvi_engine_advance(gpvi_engine *e, int nsteps);
/* 1 = outputs final, 0 = in flight, <0 = unknown id. */
int gpvi_engine_job_status(gpvi_engine *e, long long job_id);
/* Forget a finished job (frees its engine bookkeeping). */
int gpvi_engine_release(gpvi_engine *e, long long job_id);
/* Drop a job in flight: its units stop at the next step boundary; outputs undefined. */
int gpvi_engine_cancel(gpvi_engine *e, long long job_id);
int gpvi_engine_active_jobs(gpvi_engine *e);
int gpvi_engine_active_units(gpvi_engine *e);
/* Steps each active job still needs (maximum over jobs). */
int gpvi_engine_pending_steps(gpvi_engine *e);
int gpvi_engine_sync(gpvi_engine *e);

/* ---- Reference-compatible host entry points ----
 * Identical signatures to pyacofwi2D.h:2-3 (fdtd2D_modelling.c:18, :356):
 * host buffers, parameter file path; they copy to the device, run the sample
 * as a one-sample batch (gpvi_fwi2d_batch: its shots advance together; the
 * full wavefield store when it fits in 3/4 of the free device memory, else
 * checkpoints every C* = sqrt(2 nt k ntp / ni) steps) and copy back.  On
 * failure they print the message to stderr and
 * leave the outputs untouched (the reference exit()s); the status of the
 * last call on the thread is gpvi_compat_last_status(). */
void fwi_2D(char input_file[200], float *vel_inner, float *record_syn, float *record_obs, float *grad,
            int run_fwi, int verbose);
void forward_2D(char input_file[200], float *vel_inner, float *record_syn, int run_fwi, int verbose);
int gpvi_compat_last_status(void);

#ifdef __cplusplus
}
#endif
#endif /* GEOPVI_FWI2D_H */
