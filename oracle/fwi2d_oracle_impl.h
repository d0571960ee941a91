/*
 * Precision template of the CPU oracle (test infrastructure only; the
 * product never uses it).  Included twice by fwi2d_oracle.c with
 *   REAL = float,  SUF(x) = x##_f32   (reference float semantics)
 *   REAL = double, SUF(x) = x##_f64   (same scheme in double)
 *
 * Expression shapes follow fdtd2D_modelling.c so that the float build rounds
 * the way the reference's float build does: the literals 2 (int), 1.0 and
 * -2.0 (double) are kept where the reference has them, and the file is
 * compiled with -ffp-contract=off (the reference's x86-64 build has no FMA).
 */

/* Ricker wavelet, fdtd.c:1084-1093.  The squared argument is stored in a
 * float in the reference (float tmp), then the wavelet is formed in double
 * from that float and rounded to float.  The f64 twin keeps it in double. */
void SUF(ofwi_ricker)(int nt, REAL dt, REAL f0, REAL *rik)
{
    for (int it = 0; it < nt; ++it) {
        REAL arg = (REAL)it * dt - 1 / f0;
        REAL tau = (REAL)pow(OFWI_PI * f0 * arg, 2.0);
        rik[it] = (REAL)((1.0 - 2.0 * tau) * exp(-tau));
    }
}

/* Edge-replicated padding of the inner model into the absorbing frame,
 * fdtd.c:1187-1238: inner copy, then top/bottom rows of the inner columns,
 * then left/right columns over every row (which fills the corners). */
void SUF(ofwi_pad_velocity)(int pml, int ntx, int ntz, const REAL *vin, REAL *velp)
{
    const int nx = ntx - 2 * pml;
#pragma omp parallel for collapse(2)
    OFWI_XZ(ix, pml, ntx - pml, iz, pml, ntz - pml)
            velp[(long)iz * ntx + ix] = vin[(long)(iz - pml) * nx + (ix - pml)];
#pragma omp parallel for collapse(2)
    OFWI_XZ(ix, pml, ntx - pml, iz, 0, pml) {
            velp[(long)iz * ntx + ix] = velp[(long)pml * ntx + ix];
            velp[(long)(ntz - 1 - iz) * ntx + ix] = velp[(long)(ntz - pml - 1) * ntx + ix];
        }
#pragma omp parallel for collapse(2)
    for (int iz = 0; iz < ntz; ++iz)
        for (int ix = 0; ix < pml; ++ix) {
            velp[(long)iz * ntx + ix] = velp[(long)iz * ntx + pml];
            velp[(long)iz * ntx + (ntx - 1 - ix)] = velp[(long)iz * ntx + (ntx - pml - 1)];
        }
}

/* Working state of one solver call. */
typedef struct SUF(ofwi_ctx) {
    int nx, nz, ntx, ntz, pml, pml0, lc, nt, k, nr, ns, nsave;
    long ntp, ni;
    REAL dx, dz, dt;
    REAL fd[8];
    REAL *rik, *velp, *w, *t11, *t12, *t13;
    REAL *p0, *p1, *p2, *psave, *record;
    int *s_ix, s_iz, r_iz, *r_ix;
} SUF(ofwi_ctx);

/* Interior Laplacian update S1 (fdtd.c:756-780), ix-outer like the
 * reference.  GCC folds powf(x, 2) to x*x, so the square is written out. */
static void SUF(stencil)(SUF(ofwi_ctx) *c)
{
    const int L = c->lc, ntx = c->ntx, ntz = c->ntz;
    const REAL dx = c->dx, dz = c->dz, dt = c->dt;
    const REAL *fd = c->fd, *v = c->velp, *p0 = c->p0, *p1 = c->p1;
    REAL *p2 = c->p2;
#pragma omp parallel for collapse(2)
    OFWI_XZ(ix, L, ntx - L, iz, L, ntz - L) {
            const long ip = (long)iz * ntx + ix;
            REAL dxx = fd[0] * p1[ip];
            for (int m = 1; m < L; ++m)
                dxx += fd[m] * (p1[ip + m] + p1[ip - m]);
            REAL dzz = fd[0] * p1[ip];
            for (int m = 1; m < L; ++m)
                dzz += fd[m] * (p1[ip + (long)m * ntx] + p1[ip - (long)m * ntx]);
            const REAL vdt = v[ip] * dt;
            p2[ip] = 2 * p1[ip] - p0[ip] + vdt * vdt * (dxx / dx / dx + dzz / dz / dz);
        }
}

/* Second-order transmitting extrapolation along one inward direction
 * (stride st between successive inward points), fdtd.c:840-846. */
static inline REAL SUF(abc_h)(const REAL *p0, const REAL *p1, long ip, long st,
                              REAL t11, REAL t12, REAL t13)
{
    return 2 * t11 * p1[ip] + 2 * t12 * p1[ip + st] + 2 * t13 * p1[ip + 2 * st]
         - t11 * t11 * p0[ip] - 2 * t11 * t12 * p0[ip + st]
         - (t12 * t12 + 2 * t11 * t13) * p0[ip + 2 * st]
         - 2 * t12 * t13 * p0[ip + 3 * st] - t13 * t13 * p0[ip + 4 * st];
}

/* Hybrid absorbing boundary S2 (fdtd.c:833-900): four ordered sweeps,
 * top, bottom, left, right; the blend uses the double literal 1.0. */
static void SUF(abc)(SUF(ofwi_ctx) *c)
{
    const int L = c->lc, ntx = c->ntx, ntz = c->ntz, pml = c->pml;
    const REAL *p0 = c->p0, *p1 = c->p1, *w = c->w;
    const REAL *t11 = c->t11, *t12 = c->t12, *t13 = c->t13;
    REAL *p2 = c->p2;
    const long nx_l = ntx;
#pragma omp parallel for collapse(2)
    OFWI_XZ(ix, 0, ntx, iz, L, pml) {
            const long ip = (long)iz * ntx + ix;
            const REAL wt = w[iz - L];
            p2[ip] = wt * SUF(abc_h)(p0, p1, ip, nx_l, t11[ip], t12[ip], t13[ip]) + (1.0 - wt) * p2[ip];
        }
#pragma omp parallel for collapse(2)
    OFWI_XZ(ix, 0, ntx, iz, ntz - pml, ntz - L) {
            const long ip = (long)iz * ntx + ix;
            const REAL wt = w[ntz - L - iz - 1];
            p2[ip] = wt * SUF(abc_h)(p0, p1, ip, -nx_l, t11[ip], t12[ip], t13[ip]) + (1.0 - wt) * p2[ip];
        }
#pragma omp parallel for collapse(2)
    for (int iz = 0; iz < ntz; ++iz)
        for (int ix = L; ix < pml; ++ix) {
            const long ip = (long)iz * ntx + ix;
            const REAL wt = w[ix - L];
            p2[ip] = wt * SUF(abc_h)(p0, p1, ip, 1, t11[ip], t12[ip], t13[ip]) + (1.0 - wt) * p2[ip];
        }
#pragma omp parallel for collapse(2)
    for (int iz = 0; iz < ntz; ++iz)
        for (int ix = ntx - pml; ix < ntx - L; ++ix) {
            const long ip = (long)iz * ntx + ix;
            const REAL wt = w[ntx - L - ix - 1];
            p2[ip] = wt * SUF(abc_h)(p0, p1, ip, -1, t11[ip], t12[ip], t13[ip]) + (1.0 - wt) * p2[ip];
        }
}

/* Copy rotation p0 <- p1 <- p2 (fdtd.c:1053-1066); keeps p2 == p1. */
static void SUF(rotate)(SUF(ofwi_ctx) *c)
{
    const int ntx = c->ntx, ntz = c->ntz;
    REAL *p0 = c->p0, *p1 = c->p1;
    const REAL *p2 = c->p2;
#pragma omp parallel for collapse(2)
    OFWI_XZ(ix, 0, ntx, iz, 0, ntz) {
            const long ip = (long)iz * ntx + ix;
            p0[ip] = p1[ip];
            p1[ip] = p2[ip];
        }
}

static void SUF(zero_fields)(SUF(ofwi_ctx) *c)
{
    memset(c->p0, 0, sizeof(REAL) * c->ntp);
    memset(c->p1, 0, sizeof(REAL) * c->ntp);
    memset(c->p2, 0, sizeof(REAL) * c->ntp);
}

/* Forward I/O S3 (fdtd.c:902-939): source after the ABC, wavefield store
 * every k steps (second time derivative when run_fwi, else the field), and
 * the receiver line. */
static void SUF(forward_io)(SUF(ofwi_ctx) *c, int it, int is, int run_fwi)
{
    const int ntx = c->ntx, ntz = c->ntz;
    const REAL dt = c->dt;
    const long s = (long)c->s_iz * ntx + c->s_ix[is];
    c->p2[s] += c->rik[it] * dt * dt * c->velp[s] * c->velp[s];
    if (it % c->k == 0) {
        REAL *dst = c->psave + (long)(it / c->k) * c->ntp;
        const REAL *p0 = c->p0, *p1 = c->p1, *p2 = c->p2;
#pragma omp parallel for collapse(2)
        OFWI_XZ(ix, 0, ntx, iz, 0, ntz) {
                const long ip = (long)iz * ntx + ix;
                if (run_fwi)
                    dst[ip] = (p2[ip] - 2 * p1[ip] + p0[ip]) / dt / dt;
                else
                    dst[ip] = p2[ip];
            }
    }
    for (int ir = 0; ir < c->nr; ++ir)
        c->record[(long)it * c->nr + ir] = c->p2[(long)c->r_iz * ntx + c->r_ix[ir]];
}

/* Adjoint I/O S5 (fdtd.c:941-978): residual injection at every receiver,
 * then the imaging condition over the inner grid every k steps (a double
 * expression through the -2.0 literal, rounded to float on accumulation). */
static void SUF(backward_io)(SUF(ofwi_ctx) *c, int it, REAL *grad, int run_fwi)
{
    const int ntx = c->ntx, ntz = c->ntz, pml = c->pml, kk = c->k;
    const REAL dt = c->dt;
    for (int ir = 0; ir < c->nr; ++ir) {
        const long r = (long)c->r_iz * ntx + c->r_ix[ir];
        c->p2[r] += c->record[(long)it * c->nr + ir] * dt * dt * c->velp[r] * c->velp[r];
    }
    if (it % kk == 0) {
        const REAL *src = c->psave + (long)(it / kk) * c->ntp;
        const REAL *p2 = c->p2, *v = c->velp;
        const int nxi = ntx - 2 * pml;
#pragma omp parallel for collapse(2)
        OFWI_XZ(ix, pml, ntx - pml, iz, pml, ntz - pml) {
                const long ip = (long)iz * ntx + ix;
                const long gi = (long)(iz - pml) * nxi + (ix - pml);
                if (run_fwi)
                    grad[gi] += -2.0 / OFWI_POW3(v[ip]) * src[ip] * p2[ip] * kk * dt;
                else
                    grad[gi] += src[ip] * p2[ip] * kk * dt;
            }
    }
}

static void SUF(ctx_free)(SUF(ofwi_ctx) *c)
{
    free(c->rik); free(c->velp); free(c->w);
    free(c->t11); free(c->t12); free(c->t13);
    free(c->p0); free(c->p1); free(c->p2);
    free(c->psave); free(c->record); free(c->s_ix); free(c->r_ix);
}

/* Per-call setup (fdtd.c:67-240): derived sizes, FD weights, wavelet,
 * geometry, padded model and ABC coefficient arrays. */
static int SUF(ctx_init)(SUF(ofwi_ctx) *c, const ofwi_params *p, const REAL *vel_inner)
{
    int rc = ofwi_validate(p);
    if (rc) return rc;
    memset(c, 0, sizeof(*c));
    c->lc = p->Lc + 1;
    float fdf[8];
    ofwi_fd_coefficient(c->lc, fdf);
    for (int m = 0; m < c->lc; ++m) c->fd[m] = fdf[m];
    c->nx = p->nx; c->nz = p->nz; c->pml0 = p->pml0;
    c->pml = p->pml0 + c->lc;
    c->ntz = p->nz + 2 * c->pml;
    c->ntx = p->nx + 2 * c->pml;
    c->ntp = (long)c->ntz * c->ntx;
    c->ni = (long)p->nz * p->nx;
    c->nt = p->nt; c->k = p->nt_interval; c->nr = p->nr; c->ns = p->ns;
    c->dx = p->dx; c->dz = p->dz; c->dt = p->dt;
    c->nsave = p->nt / c->k + (p->nt % c->k != 0);

    c->rik = (REAL *)malloc(sizeof(REAL) * c->nt);
    SUF(ofwi_ricker)(c->nt, (REAL)p->dt, (REAL)p->f0, c->rik);

    c->s_ix = (int *)malloc(sizeof(int) * c->ns);
    for (int is = 0; is < c->ns; ++is) c->s_ix[is] = is * p->ds + c->pml + p->ns0;
    c->s_iz = c->pml + p->depths;
    c->r_iz = c->pml + p->depthr;
    c->r_ix = (int *)malloc(sizeof(int) * c->nr);
    for (int ir = 0; ir < c->nr; ++ir) c->r_ix[ir] = c->pml + ir * p->dr + p->nr0;

    c->p0 = (REAL *)calloc(c->ntp, sizeof(REAL));
    c->p1 = (REAL *)calloc(c->ntp, sizeof(REAL));
    c->p2 = (REAL *)calloc(c->ntp, sizeof(REAL));
    c->psave = (REAL *)calloc((size_t)c->ntp * c->nsave, sizeof(REAL));
    c->record = (REAL *)calloc((size_t)c->nt * c->nr, sizeof(REAL));
    c->velp = (REAL *)malloc(sizeof(REAL) * c->ntp);
    SUF(ofwi_pad_velocity)(c->pml, c->ntx, c->ntz, vel_inner, c->velp);

    c->w = (REAL *)malloc(sizeof(REAL) * (c->pml0 > 0 ? c->pml0 : 1));
    for (int i = 0; i < c->pml0; ++i) c->w[i] = 1.0 - 1.0 * i / c->pml0;
    c->t11 = (REAL *)malloc(sizeof(REAL) * c->ntp);
    c->t12 = (REAL *)malloc(sizeof(REAL) * c->ntp);
    c->t13 = (REAL *)malloc(sizeof(REAL) * c->ntp);
    const REAL alpha = 1.0, dt = c->dt, dx = c->dx;
    const int ntx = c->ntx, ntz = c->ntz;
#pragma omp parallel for collapse(2)
    OFWI_XZ(ix, 0, ntx, iz, 0, ntz) {
            const long ip = (long)iz * ntx + ix;
            const REAL s = alpha * c->velp[ip] * dt / dx;
            c->t11[ip] = (2 - s) * (1 - s) / 2;
            c->t12[ip] = s * (2 - s);
            c->t13[ip] = s * (s - 1) / 2;
        }
    if (!c->rik || !c->p0 || !c->p1 || !c->p2 || !c->psave || !c->record || !c->velp ||
        !c->t11 || !c->t12 || !c->t13 || !c->s_ix || !c->r_ix) {
        SUF(ctx_free)(c);
        return OFWI_ENOMEM;
    }
    return 0;
}

static void SUF(propagate)(SUF(ofwi_ctx) *c)
{
    SUF(stencil)(c);
    SUF(abc)(c);
}

/* forward_aco_2D, fdtd.c:678-714 */
static void SUF(forward_pass)(SUF(ofwi_ctx) *c, int is, int run_fwi)
{
    SUF(zero_fields)(c);
    for (int it = 0; it < c->nt; ++it) {
        SUF(propagate)(c);
        SUF(forward_io)(c, it, is, run_fwi);
        SUF(rotate)(c);
    }
}

/* backward_aco_2D, fdtd.c:717-753 */
static void SUF(backward_pass)(SUF(ofwi_ctx) *c, REAL *grad, int run_fwi)
{
    SUF(zero_fields)(c);
    for (int it = c->nt - 1; it >= 0; --it) {
        SUF(propagate)(c);
        SUF(backward_io)(c, it, grad, run_fwi);
        SUF(rotate)(c);
    }
}

int SUF(ofwi_fwi)(const ofwi_params *p, const REAL *vel_inner, REAL *record_syn,
                  const REAL *record_obs, REAL *grad, int run_fwi, int shot_begin, int shot_end)
{
    SUF(ofwi_ctx) c;
    int rc = SUF(ctx_init)(&c, p, vel_inner);
    if (rc) return rc;
    if (shot_end < 0 || shot_end > c.ns) shot_end = c.ns;
    if (shot_begin < 0) shot_begin = 0;
    const long rec = (long)c.nt * c.nr;
    for (int is = shot_begin; is < shot_end; ++is) {
        SUF(forward_pass)(&c, is, run_fwi);
        /* record copy and residual, fdtd.c:269-278 */
        REAL *syn = record_syn + is * rec;
        const REAL *obs = record_obs + is * rec;
#pragma omp parallel for
        for (long i = 0; i < rec; ++i) {
            syn[i] = c.record[i];
            if (run_fwi) c.record[i] -= obs[i];
            else c.record[i] = obs[i];
        }
        SUF(backward_pass)(&c, grad, run_fwi);
    }
    SUF(ctx_free)(&c);
    return 0;
}

int SUF(ofwi_forward)(const ofwi_params *p, const REAL *vel_inner, REAL *record_syn)
{
    SUF(ofwi_ctx) c;
    int rc = SUF(ctx_init)(&c, p, vel_inner);
    if (rc) return rc;
    const long rec = (long)c.nt * c.nr;
    for (int is = 0; is < c.ns; ++is) {
        SUF(forward_pass)(&c, is, 1);
        memcpy(record_syn + is * rec, c.record, sizeof(REAL) * rec);
    }
    SUF(ctx_free)(&c);
    return 0;
}

void SUF(ofwi_step)(const ofwi_params *p, const REAL *velp, const REAL *p0, const REAL *p1, REAL *p2)
{
    /* A light context around caller-owned fields. */
    SUF(ofwi_ctx) c;
    memset(&c, 0, sizeof(c));
    c.lc = p->Lc + 1;
    float fdf[8];
    ofwi_fd_coefficient(c.lc, fdf);
    for (int m = 0; m < c.lc; ++m) c.fd[m] = fdf[m];
    c.pml0 = p->pml0;
    c.pml = p->pml0 + c.lc;
    c.ntz = p->nz + 2 * c.pml;
    c.ntx = p->nx + 2 * c.pml;
    c.ntp = (long)c.ntz * c.ntx;
    c.dx = p->dx; c.dz = p->dz; c.dt = p->dt;
    c.velp = (REAL *)velp;
    c.p0 = (REAL *)p0; c.p1 = (REAL *)p1; c.p2 = p2;
    REAL wbuf[256];
    c.w = wbuf;
    for (int i = 0; i < c.pml0 && i < 256; ++i) c.w[i] = 1.0 - 1.0 * i / c.pml0;
    c.t11 = (REAL *)malloc(sizeof(REAL) * c.ntp);
    c.t12 = (REAL *)malloc(sizeof(REAL) * c.ntp);
    c.t13 = (REAL *)malloc(sizeof(REAL) * c.ntp);
    const REAL alpha = 1.0, dt = c.dt, dx = c.dx;
    for (long ip = 0; ip < c.ntp; ++ip) {
        const REAL s = alpha * velp[ip] * dt / dx;
        c.t11[ip] = (2 - s) * (1 - s) / 2;
        c.t12[ip] = s * (2 - s);
        c.t13[ip] = s * (s - 1) / 2;
    }
    memcpy(p2, p1, sizeof(REAL) * c.ntp);
    SUF(propagate)(&c);
    free(c.t11); free(c.t12); free(c.t13);
}
