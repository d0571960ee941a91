/*
 * CPU oracle for the 2-D acoustic FDTD misfit/gradient path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against; nothing in the product (geopvi_amd/)
 * links, loads or calls it.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it.
 *
 * It restates geopvi/forward/fwi2d/fdtd2D_modelling.c (finite-difference
 * path only; the FFTW pseudo-spectral option is not restated) in plain C
 * with OpenMP, keeping the reference's loop order, expression typing
 * (float arithmetic with the double-typed literals where the reference has
 * them), copy-based field rotation, full wavefield store and per-call setup,
 * so that it is both a numerical checker and a faithful CPU timing baseline.
 *
 * Two precisions are built from one template: *_f32 follows the reference's
 * float semantics, *_f64 evaluates the same discrete scheme in double.
 */
#ifndef GEOPVI_FWI2D_ORACLE_H
#define GEOPVI_FWI2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* The 19 run parameters, in the order the reference's parameter file lists
 * them (fdtd2D_modelling.c:1337-1356).  Lc is the configured half-order
 * (the solver uses Lc+1). */
typedef struct ofwi_params {
    int nx, nz, pml0, Lc, laplace_slover, ns, nt, ds, ns0;
    int depths, depthr, nr, dr, nr0, nt_interval;
    float dx, dz, dt, f0;
} ofwi_params;

/* Parse the positional parameter file (fdtd2D_modelling.c:1328-1379).
 * Returns 0 on success, -1 if the file cannot be opened, -2 on a parse error. */
int ofwi_read_params(const char *path, ofwi_params *p);

/* Geometry/parameter validation; 0 if the FD path can run, else <0. */
int ofwi_validate(const ofwi_params *p);

void ofwi_set_threads(int n);
int ofwi_get_threads(void);

/* Setup pieces, exposed for unit checks. */
int ofwi_fd_coefficient(int lc_solver, float *fd);                 /* fdtd.c:1272-1326 */
void ofwi_ricker_f32(int nt, float dt, float f0, float *rik);       /* fdtd.c:1084-1093 */
void ofwi_ricker_f64(int nt, double dt, double f0, double *rik);
void ofwi_pad_velocity_f32(int pml, int ntx, int ntz, const float *vel_inner, float *velp);   /* fdtd.c:1187-1238 */
void ofwi_pad_velocity_f64(int pml, int ntx, int ntz, const double *vel_inner, double *velp);

/* One propagation step without I/O: stencil + ordered ABC sweeps on padded
 * fields (fdtd.c:756-780, 833-900).  p2 is overwritten with p1 first
 * (the copy rotation leaves p2 == p1 at the start of every step). */
void ofwi_step_f32(const ofwi_params *p, const float *velp, const float *p0, const float *p1, float *p2);
void ofwi_step_f64(const ofwi_params *p, const double *velp, const double *p0, const double *p1, double *p2);

/* Full misfit-gradient evaluation for one velocity sample (fdtd.c:18-353),
 * over shots [shot_begin, shot_end) (pass 0, -1 for all shots).  record_syn
 * and record_obs are [ns][nt][nr]; grad is [nz][nx] and is accumulated (+=).
 * Returns 0 on success, <0 on invalid parameters. */
int ofwi_fwi_f32(const ofwi_params *p, const float *vel_inner, float *record_syn,
                 const float *record_obs, float *grad, int run_fwi, int shot_begin, int shot_end);
int ofwi_fwi_f64(const ofwi_params *p, const double *vel_inner, double *record_syn,
                 const double *record_obs, double *grad, int run_fwi, int shot_begin, int shot_end);

/* Forward modelling only (fdtd.c:356-675). */
int ofwi_forward_f32(const ofwi_params *p, const float *vel_inner, float *record_syn);
int ofwi_forward_f64(const ofwi_params *p, const double *vel_inner, double *record_syn);

#ifdef __cplusplus
}
#endif
#endif
