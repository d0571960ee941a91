/*
 * CPU oracle for the 2-D acoustic FDTD misfit/gradient path.
 * TEST INFRASTRUCTURE ONLY: used by tests/, smoke() and the CPU baseline of
 * bench.py as the checker; the product path never links or calls it.
 *
 * Restates geopvi/forward/fwi2d/fdtd2D_modelling.c (FD path): see
 * fwi2d_oracle.h for what is kept from the reference and why.
 */
#include "fwi2d_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define OFWI_PI 3.1415926 /* the reference's own value of pi, fdtd.c:15 */
#define OFWI_ENOMEM (-10)

/* Positional parameter file, fdtd.c:1328-1379: every value is preceded by
 * exactly one comment line; 15 integers then 4 floats.  Whitespace (blank
 * lines included) after a value is skipped before the next comment line. */
static int read_one(FILE *fp, const char *fmt, void *dst)
{
    char line[256];
    if (!fgets(line, sizeof line, fp)) return -2;
    if (fscanf(fp, fmt, dst) != 1) return -2;
    if (fscanf(fp, " ") < 0 && !feof(fp)) return -2;
    return 0;
}

int ofwi_read_params(const char *path, ofwi_params *p)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    int *ints[15] = {&p->nx, &p->nz, &p->pml0, &p->Lc, &p->laplace_slover, &p->ns, &p->nt,
                     &p->ds, &p->ns0, &p->depths, &p->depthr, &p->nr, &p->dr, &p->nr0,
                     &p->nt_interval};
    float *flts[4] = {&p->dx, &p->dz, &p->dt, &p->f0};
    int rc = 0;
    for (int i = 0; i < 15 && !rc; ++i) rc = read_one(fp, " %d", ints[i]);
    for (int i = 0; i < 4 && !rc; ++i) rc = read_one(fp, " %f", flts[i]);
    fclose(fp);
    return rc;
}

int ofwi_validate(const ofwi_params *p)
{
    if (p->laplace_slover != 0) return -3;              /* pseudo-spectral path not restated */
    if (p->Lc < 3 || p->Lc > 6) return -4;              /* fd_coefficient covers 3..6 */
    if (p->nx < 1 || p->nz < 1 || p->pml0 < 0 || p->ns < 1 || p->nt < 1 || p->nr < 1 ||
        p->nt_interval < 1)
        return -5;
    if (!(p->dx > 0) || !(p->dz > 0) || !(p->dt > 0)) return -5;
    if (p->nr > 1 && p->dr < 1) return -6;              /* coincident receivers */
    const int lc = p->Lc + 1, pml = p->pml0 + lc;
    const int ntx = p->nx + 2 * pml, ntz = p->nz + 2 * pml;
    const long s_first = (long)pml + p->ns0, s_last = s_first + (long)(p->ns - 1) * p->ds;
    const long r_first = (long)pml + p->nr0, r_last = r_first + (long)(p->nr - 1) * p->dr;
    const long s_iz = (long)pml + p->depths, r_iz = (long)pml + p->depthr;
    if (s_first < 0 || s_last < 0 || s_first >= ntx || s_last >= ntx) return -7;
    if (r_first < 0 || r_last < 0 || r_first >= ntx || r_last >= ntx) return -7;
    if (s_iz < 0 || s_iz >= ntz || r_iz < 0 || r_iz >= ntz) return -7;
    return 0;
}

void ofwi_set_threads(int n)
{
    if (n > 0) omp_set_num_threads(n);
}

int ofwi_get_threads(void) { return omp_get_max_threads(); }

/* Central second-derivative weights for solver half-width lc (= Lc+1),
 * fdtd.c:1272-1326, written as exact rationals (Fornberg's closed form). */
int ofwi_fd_coefficient(int lc, float *fd)
{
    switch (lc) {
    case 4:
        fd[0] = -49.0 / 18; fd[1] = 3.0 / 2; fd[2] = -3.0 / 20; fd[3] = 1.0 / 90;
        return 0;
    case 5:
        fd[0] = -205.0 / 72; fd[1] = 8.0 / 5; fd[2] = -1.0 / 5; fd[3] = 8.0 / 315;
        fd[4] = -1.0 / 560;
        return 0;
    case 6:
        fd[0] = -5269.0 / 1800; fd[1] = 5.0 / 3; fd[2] = -5.0 / 21; fd[3] = 5.0 / 126;
        fd[4] = -5.0 / 1008; fd[5] = 1.0 / 3150;
        return 0;
    case 7:
        fd[0] = -5369.0 / 1800; fd[1] = 12.0 / 7; fd[2] = -15.0 / 56; fd[3] = 10.0 / 189;
        fd[4] = -1.0 / 112; fd[5] = 2.0 / 1925; fd[6] = -1.0 / 16632;
        return 0;
    default:
        return -4;
    }
}

/* Loop order of the grid sweeps.  The reference walks its row-major arrays
 * ix-outer (fdtd.c:768 and every sweep after it), and the default build keeps
 * that order so that it is a faithful CPU timing baseline.  Every point of a
 * sweep is computed from values the sweep does not write, so the order
 * changes no result: OFWI_ROW_MAJOR builds the same arithmetic iz-outer
 * (libfwi2d_oracle_rm.so, a faster checker for long fixture runs; the tests
 * hold the two builds to identical bits). */
#ifdef OFWI_ROW_MAJOR
#define OFWI_XZ(ix, x0, x1, iz, z0, z1) \
    for (int iz = (z0); iz < (z1); ++iz) for (int ix = (x0); ix < (x1); ++ix)
#else
#define OFWI_XZ(ix, x0, x1, iz, z0, z1) \
    for (int ix = (x0); ix < (x1); ++ix) for (int iz = (z0); iz < (z1); ++iz)
#endif

/* float build: reference semantics */
#define REAL float
#define SUF(x) x##_f32
#define OFWI_POW3(v) powf((v), 3)
#include "fwi2d_oracle_impl.h"
#undef REAL
#undef SUF
#undef OFWI_POW3

/* double twin */
#define REAL double
#define SUF(x) x##_f64
#define OFWI_POW3(v) pow((v), 3)
#include "fwi2d_oracle_impl.h"
#undef REAL
#undef SUF
#undef OFWI_POW3
