// CPU run of the pseudo-spectral Laplacian's transform phases
// (geopvi_amd/csrc/fwi2d_spectral.h): every phase that the device runs
// between two barriers runs here for each thread of the workgroup in turn,
// over the same workgroup tiling (row pairs per group, spectrum columns per
// group), so the index arithmetic the kernels use is checked on the CPU
// against numpy (tests/test_spectral_cpu.py).  Test infrastructure only.
#define GPVI_SPEC_FN inline
#include "fwi2d_spectral.h"

#include <cmath>
#include <utility>
#include <vector>

using namespace gpvi::spec;

namespace {

std::vector<Cx> table(int n)
{
    std::vector<Cx> t(n);
    for (int m = 0; m < n; ++m) {
        const double a = 2.0 * M_PI * (double)m / (double)n;
        t[m] = cx(std::cos(a), std::sin(a));
    }
    return t;
}

// The plan's passes, X -> Y -> X ...; returns the buffer holding the result.
Cx* transform(Cx* x, Cx* y, int nseq, const Plan& p, const Cx* t, int sgn, int nthr)
{
    int ns = 1;
    for (int i = 0; i < p.npass; ++i) {
        for (int tid = 0; tid < nthr; ++tid) run_pass<13>(p.radix[i], x, y, nseq, p.n, ns, t, sgn, tid, nthr);
        ns *= p.radix[i];
        std::swap(x, y);
    }
    return x;
}

}  // namespace

extern "C" {

int spec_plan(int n, int* radix)
{
    Plan p;
    if (!make_plan(n, &p)) return -1;
    for (int i = 0; i < p.npass; ++i) radix[i] = p.radix[i];
    return p.npass;
}

int spec_pairs_per_group(int ntx) { return pairs_per_group(ntx); }
int spec_cols_per_group(int ntz) { return cols_per_group(ntz); }

// nseq transforms of length n (interleaved re/im in and out).
int spec_fft(const double* in, int n, int nseq, int sgn, int nthr, double* out)
{
    Plan p;
    if (!make_plan(n, &p)) return -1;
    std::vector<Cx> x(nseq * n), y(nseq * n);
    for (int i = 0; i < nseq * n; ++i) x[i] = cx(in[2 * i], in[2 * i + 1]);
    const std::vector<Cx> t = table(n);
    const Cx* r = transform(x.data(), y.data(), nseq, p, t.data(), sgn, nthr);
    for (int i = 0; i < nseq * n; ++i) {
        out[2 * i] = r[i].re;
        out[2 * i + 1] = r[i].im;
    }
    return 0;
}

// Re(IDFT2(-k * DFT2(p1))), unnormalised, of one ntz x ntx field through the
// three phases: np row pairs and ncol spectrum columns per workgroup of nthr
// threads.
int spec_laplacian(const float* p1, const float* k, int ntz, int ntx, int np, int ncol, int nthr, double* out)
{
    Plan px, pz;
    if (!make_plan(ntx, &px) || !make_plan(ntz, &pz) || np < 1 || ncol < 1) return -1;
    const int nc = ntx / 2 + 1;
    const int groups = ((ntz + 1) / 2 + np - 1) / np;
    std::vector<Cx> s((size_t)ntz * nc, cx(NAN, NAN));
    const std::vector<Cx> tx = table(ntx), tz = table(ntz);

    for (int g = 0; g < groups; ++g) {
        std::vector<Cx> x(np * ntx), y(np * ntx);
        const int pair0 = g * np;
        for (int tid = 0; tid < nthr; ++tid)
            rows_load(x.data(), np, ntx, ntz, pair0, [&](int iz, int ix) { return (double)p1[iz * ntx + ix]; },
                      tid, nthr);
        const Cx* z = transform(x.data(), y.data(), np, px, tx.data(), +1, nthr);
        for (int tid = 0; tid < nthr; ++tid)
            rows_unpack(z, np, ntx, ntz, nc, pair0, [&](int iz, int f, Cx v) { s[(size_t)iz * nc + f] = v; }, tid,
                        nthr);
    }
    for (int f0 = 0; f0 < nc; f0 += ncol) {
        std::vector<Cx> x(ncol * ntz), y(ncol * ntz);
        for (int tid = 0; tid < nthr; ++tid)
            cols_load(x.data(), ncol, ntz, nc, f0, [&](int iz, int f) { return s[(size_t)iz * nc + f]; }, tid, nthr);
        Cx* z = transform(x.data(), y.data(), ncol, pz, tz.data(), +1, nthr);
        Cx* other = z == x.data() ? y.data() : x.data();
        for (int tid = 0; tid < nthr; ++tid)
            cols_scale(z, ncol, ntz, nc, f0, [&](int iz, int f) { return k[iz * ntx + f]; }, tid, nthr);
        const Cx* w = transform(z, other, ncol, pz, tz.data(), -1, nthr);
        for (int tid = 0; tid < nthr; ++tid)
            cols_store(w, ncol, ntz, nc, f0, [&](int iz, int f, Cx v) { s[(size_t)iz * nc + f] = v; }, tid, nthr);
    }
    for (int g = 0; g < groups; ++g) {
        std::vector<Cx> x(np * ntx), y(np * ntx);
        const int pair0 = g * np;
        for (int tid = 0; tid < nthr; ++tid)
            rows_repack(x.data(), np, ntx, ntz, pair0, [&](int iz, int f) { return s[(size_t)iz * nc + f]; }, tid,
                        nthr);
        const Cx* z = transform(x.data(), y.data(), np, px, tx.data(), -1, nthr);
        for (int q = 0; q < np; ++q)
            for (int w = 0; w < 2; ++w) {
                const int iz = 2 * (pair0 + q) + w;
                if (iz >= ntz) continue;
                for (int ix = 0; ix < ntx; ++ix)
                    out[iz * ntx + ix] = w ? z[q * ntx + ix].im : z[q * ntx + ix].re;
            }
    }
    return 0;
}

}  // extern "C"
