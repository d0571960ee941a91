// Batched mixed-radix discrete Fourier transforms held in LDS, for the
// pseudo-spectral Laplacian of the FWI-2D solver (fwi2d_psm.hip).
//
// The reference (geopvi/forward/fwi2d/fdtd2D_modelling.c:783-830) applies,
// every time step, Re(IDFT2(-k * DFT2(p1))) with FFTW's unscaled forward
// (e^{-i...}) and backward (e^{+i...}) 2-D transforms and the float table
// k = |k|^2 of wavenumber (fdtd.c:1127-1184).  The same operator is computed
// here in three passes over a per-unit half spectrum S (ntz rows of
// nc = ntx/2 + 1 complex doubles):
//
//   rows forward    two real rows a, b of p1 form z = a + i b; one complex
//                   transform of length ntx gives Z = A + i B, and the half
//                   spectra A[f] = (Z[f] + conj Z[-f]) / 2 and
//                   B[f] = (Z[f] - conj Z[-f]) / 2i of the two rows go to S
//   columns         each spectrum column: forward transform of length ntz,
//                   multiplication by -k[kz][kx] (the reference's float k),
//                   backward transform
//   rows backward   two rows of S, extended by Hermitian symmetry
//                   (S[iz][ntx - f] = conj S[iz][f]; the imaginary parts of
//                   the f = 0 and f = ntx/2 terms are dropped, as a
//                   complex-to-real transform does), form Z = A + i B; one
//                   complex backward transform gives both rows' real results
//                   (real and imaginary part)
//
// The scaled 2-D spectrum is Hermitian (k is even in both axes), so its
// backward transform is real and equals the reference's real part; the
// values differ from FFTW's only by the rounding of the transforms, as
// between any two FFT libraries.  Every transform is a Stockham
// (self-sorting) sequence of radix-R passes, R in {4, 2, 3, 5, 7, 11, 13},
// between two LDS buffers; the odd radices use the symmetric form of the
// length-R DFT (pairs q, R - q share their cosine and sine sums).
//
// This header holds the index arithmetic of every phase as plain functions of
// (thread, thread count) so that it compiles for the device (the includer
// defines nothing) and for a CPU check (the includer defines GPVI_SPEC_FN as
// `inline` and runs each phase for every thread in turn, which is what a
// barrier between phases guarantees on the device).
#pragma once

#ifndef GPVI_SPEC_FN
#define GPVI_SPEC_FN __host__ __device__ inline
#endif

namespace gpvi {
namespace spec {

struct alignas(16) Cx {
    double re, im;
};

GPVI_SPEC_FN Cx cx(double r, double i)
{
    Cx c;
    c.re = r;
    c.im = i;
    return c;
}
GPVI_SPEC_FN Cx add(Cx a, Cx b) { return cx(a.re + b.re, a.im + b.im); }
GPVI_SPEC_FN Cx sub(Cx a, Cx b) { return cx(a.re - b.re, a.im - b.im); }
GPVI_SPEC_FN Cx mul(Cx a, Cx b) { return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
// -i * sgn * x
GPVI_SPEC_FN Cx rot(Cx x, int sgn) { return sgn > 0 ? cx(x.im, -x.re) : cx(-x.im, x.re); }

constexpr int kMaxPass = 16;
constexpr int kThreads = 256;        // workgroup of the transform kernels
constexpr int kLdsCx = 4096;         // LDS budget of one workgroup, in Cx (64 KiB)
// Row pairs and spectrum columns per workgroup, at most: on the 200 x 400
// model with 10-point bands (220 x 420) one step of 4 samples x 20 shots took
// 150.7 us with (2, 4), 161.1 with (2, 8), 169.1 with (1, 4), 169.3 with
// (2, 2), 179.2 with (4, 4), 193.0 with (1, 2) (256 threads; 128 threads were
// slower in an earlier sweep).
constexpr int kMaxPairs = 2;
constexpr int kMaxCols = 4;

// The radices of a length-n transform, in pass order.
struct Plan {
    int n, npass;
    int radix[kMaxPass];
};

// Factor n into passes of radix 4 (as many as fit), then one 2 if needed, then
// 3, 5, 7, 11, 13.  False when n < 1 or has another prime factor.
inline bool make_plan(int n, Plan* p)
{
    p->n = n;
    p->npass = 0;
    if (n < 1) return false;
    int m = n;
    while (m % 4 == 0) {
        if (p->npass == kMaxPass) return false;
        p->radix[p->npass++] = 4;
        m /= 4;
    }
    static const int odd[] = {2, 3, 5, 7, 11, 13};
    for (int r : odd) {
        while (m % r == 0) {
            if (p->npass == kMaxPass) return false;
            p->radix[p->npass++] = r;
            m /= r;
        }
    }
    return m == 1;
}

// Twiddle table of a length-n transform in LDS: (cos, sin)(2 pi m / n) for
// m < n / 2 when n is even (the second half is the first negated, exactly),
// else for m < n.
GPVI_SPEC_FN int table_len(int n) { return (n & 1) == 0 ? n >> 1 : n; }
GPVI_SPEC_FN Cx twiddle(const Cx* T, int m, int n)
{
    const int h = (n & 1) == 0 ? n >> 1 : n;
    const bool neg = m >= h;
    const Cx w = T[neg ? m - h : m];
    return neg ? cx(-w.re, -w.im) : w;
}

// Row pairs per workgroup for rows of length ntx, spectrum columns per
// workgroup for columns of length ntz (twiddle table + two buffers within
// the LDS budget); 0 when one does not fit.
inline int pairs_per_group(int ntx)
{
    const int r = (kLdsCx - table_len(ntx)) / (2 * ntx);
    return r < 1 ? 0 : (r < kMaxPairs ? r : kMaxPairs);
}
inline int cols_per_group(int ntz)
{
    const int c = (kLdsCx - table_len(ntz)) / (2 * ntz);
    return c < 1 ? 0 : (c < kMaxCols ? c : kMaxCols);
}

// The length-R DFT of v, W = e^{-sgn 2 pi i / R}: d[q * ns] = sum_r v[r] W^{rq}.
// cs[m], sn[m] = cos, sin(2 pi m / R), m = 1..(R-1)/2.  v is overwritten
// (the odd radices form the sums and differences of the pairs r, R - r in
// place and store each output pair as soon as it is made, so that no second
// copy of the R values is live).
template <int R>
GPVI_SPEC_FN void dft_store(Cx* v, const double* cs, const double* sn, int sgn, Cx* d, int ns)
{
    if constexpr (R == 2) {
        d[0] = add(v[0], v[1]);
        d[ns] = sub(v[0], v[1]);
    } else if constexpr (R == 4) {
        const Cx t0 = add(v[0], v[2]), t1 = sub(v[0], v[2]);
        const Cx t2 = add(v[1], v[3]), t3 = rot(sub(v[1], v[3]), sgn);
        d[0] = add(t0, t2);
        d[2 * ns] = sub(t0, t2);
        d[ns] = add(t1, t3);
        d[3 * ns] = sub(t1, t3);
    } else {
        constexpr int H = (R - 1) / 2;
        Cx sum = v[0];
#pragma unroll
        for (int r = 1; r <= H; ++r) {
            const Cx a = v[r], b = v[R - r];
            v[r] = add(a, b);       // s_r
            v[R - r] = sub(a, b);   // d_r
            sum = add(sum, v[r]);
        }
        d[0] = sum;
#pragma unroll
        for (int q = 1; q <= H; ++q) {
            Cx a = v[0], b = cx(0.0, 0.0);
#pragma unroll
            for (int r = 1; r <= H; ++r) {
                const int m = (r * q) % R;
                const double c = m <= H ? cs[m] : cs[R - m];
                const double si = m <= H ? sn[m] : -sn[R - m];
                a = cx(a.re + c * v[r].re, a.im + c * v[r].im);
                b = cx(b.re + si * v[R - r].re, b.im + si * v[R - r].im);
            }
            const Cx rb = rot(b, sgn);
            d[q * ns] = add(a, rb);
            d[(R - q) * ns] = sub(a, rb);
        }
    }
}

// One Stockham pass of radix R over nseq sequences of length n stored back to
// back (src -> dst): the butterflies b = tid, tid + nthr, ... of nseq * n / R.
// ns = product of the radices of the earlier passes.  T: the twiddle table
// (table_len, twiddle).  sgn = +1: the e^{-i} transform, -1: e^{+i}.
template <int R>
GPVI_SPEC_FN void pass(const Cx* src, Cx* dst, int nseq, int n, int ns, const Cx* T, int sgn, int tid, int nthr)
{
    constexpr int H = (R - 1) / 2;
    const int nr = n / R;
    const int tstep = n / (ns * R);
    double cs[H + 1], sn[H + 1];
    cs[0] = 1.0;
    sn[0] = 0.0;
    if constexpr ((R & 1) != 0) {
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            const Cx w = twiddle(T, m * nr, n);
            cs[m] = w.re;
            sn[m] = w.im;
        }
    }
    for (int b = tid; b < nseq * nr; b += nthr) {
        const int q = b / nr, j = b - q * nr;
        const Cx* s = src + q * n;
        Cx* d = dst + q * n;
        const int k = j % ns;
        Cx v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = s[j + r * nr];
        if (k > 0) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                const Cx w = twiddle(T, r * k * tstep, n);
                v[r] = mul(v[r], cx(w.re, -sgn * w.im));
            }
        }
        dft_store<R>(v, cs, sn, sgn, d + (j - k) * R + k, ns);
    }
}

// The pass of radix r (one of the plan's radices, all <= MaxR: a kernel that
// needs only the small radices is not held to the registers of the large).
template <int MaxR>
GPVI_SPEC_FN void run_pass(int r, const Cx* src, Cx* dst, int nseq, int n, int ns, const Cx* T, int sgn, int tid,
                           int nthr)
{
    if (r == 4) {
        pass<4>(src, dst, nseq, n, ns, T, sgn, tid, nthr);
    } else if (r == 2) {
        pass<2>(src, dst, nseq, n, ns, T, sgn, tid, nthr);
    } else if (r == 3) {
        pass<3>(src, dst, nseq, n, ns, T, sgn, tid, nthr);
    } else if (r == 5) {
        pass<5>(src, dst, nseq, n, ns, T, sgn, tid, nthr);
    } else if (r == 7) {
        pass<7>(src, dst, nseq, n, ns, T, sgn, tid, nthr);
    } else if constexpr (MaxR >= 11) {
        if (r == 11) {
            pass<11>(src, dst, nseq, n, ns, T, sgn, tid, nthr);
        } else if constexpr (MaxR >= 13) {
            pass<13>(src, dst, nseq, n, ns, T, sgn, tid, nthr);
        }
    }
}

// The largest radix of a plan.
inline int max_radix(const Plan& p)
{
    int m = 1;
    for (int i = 0; i < p.npass; ++i) m = p.radix[i] > m ? p.radix[i] : m;
    return m;
}

// for t = tid, tid + nthr, ... < n: put(t, get(t)), kChunk elements at a
// time with every get of a chunk before its puts, so that a thread's global
// loads are in flight together instead of one latency per element.
constexpr int kChunk = 4;
template <class Get, class Put>
GPVI_SPEC_FN void chunked(int n, int tid, int nthr, Get get, Put put)
{
    using V = decltype(get(0));
    for (int t0 = tid; t0 < n; t0 += kChunk * nthr) {
        V v[kChunk];
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            const int t = t0 + j * nthr;
            if (t < n) v[j] = get(t);
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            const int t = t0 + j * nthr;
            if (t < n) put(t, v[j]);
        }
    }
}

// Rows forward, load: sequence q of X holds real rows 2 (pair0 + q) (real
// part) and 2 (pair0 + q) + 1 (imaginary part) of p1; rows past ntz are 0.
template <class Load>
GPVI_SPEC_FN void rows_load(Cx* X, int np, int ntx, int ntz, int pair0, Load p1, int tid, int nthr)
{
    chunked(
        np * ntx, tid, nthr,
        [&](int t) {
            const int q = t / ntx, ix = t - q * ntx;
            const int iz = 2 * (pair0 + q);
            const double a = iz < ntz ? p1(iz, ix) : 0.0;
            const double b = iz + 1 < ntz ? p1(iz + 1, ix) : 0.0;
            return cx(a, b);
        },
        [&](int t, Cx v) { X[t] = v; });
}

// Rows forward, unpack: the half spectra f = 0..nc-1 of both rows of each
// pair from the transform Z of a + i b; st(iz, f, value) for rows < ntz.
template <class Store>
GPVI_SPEC_FN void rows_unpack(const Cx* Z, int np, int ntx, int ntz, int nc, int pair0, Store st, int tid, int nthr)
{
    for (int t = tid; t < np * nc; t += nthr) {
        const int q = t / nc, f = t - q * nc;
        const int iz = 2 * (pair0 + q);
        if (iz >= ntz) continue;
        const Cx zf = Z[q * ntx + f];
        const Cx zn = Z[q * ntx + (f == 0 ? 0 : ntx - f)];
        st(iz, f, cx(0.5 * (zf.re + zn.re), 0.5 * (zf.im - zn.im)));
        if (iz + 1 < ntz) st(iz + 1, f, cx(0.5 * (zf.im + zn.im), -0.5 * (zf.re - zn.re)));
    }
}

// Columns, load: sequence c of X holds spectrum column f0 + c (all ntz rows);
// columns past nc are 0.
template <class Load>
GPVI_SPEC_FN void cols_load(Cx* X, int nc_tile, int ntz, int nc, int f0, Load sp, int tid, int nthr)
{
    chunked(
        nc_tile * ntz, tid, nthr,
        [&](int t) {
            const int iz = t / nc_tile, c = t - iz * nc_tile;
            const int f = f0 + c;
            return f < nc ? sp(iz, f) : cx(0.0, 0.0);
        },
        [&](int t, Cx v) {
            const int iz = t / nc_tile, c = t - iz * nc_tile;
            X[c * ntz + iz] = v;
        });
}

// Columns, scale: X[c][kz] *= -k(kz, f0 + c) (the float k widened).
template <class K>
GPVI_SPEC_FN void cols_scale(Cx* X, int nc_tile, int ntz, int nc, int f0, K k, int tid, int nthr)
{
    for (int t = tid; t < nc_tile * ntz; t += nthr) {
        const int c = t / ntz, iz = t - c * ntz;
        const int f = f0 + c;
        if (f >= nc) continue;
        const double m = (double)(-k(iz, f));
        X[t] = cx(m * X[t].re, m * X[t].im);
    }
}

// Columns, store: st(iz, f, value) for the tile's columns < nc.
template <class Store>
GPVI_SPEC_FN void cols_store(const Cx* X, int nc_tile, int ntz, int nc, int f0, Store st, int tid, int nthr)
{
    for (int t = tid; t < nc_tile * ntz; t += nthr) {
        const int iz = t / nc_tile, c = t - iz * nc_tile;
        const int f = f0 + c;
        if (f < nc) st(iz, f, X[c * ntz + iz]);
    }
}

// Rows backward, load: sequence q of X = A + i B over the full length ntx,
// where A, B are the Hermitian extensions of the half-spectrum rows
// 2 (pair0 + q) and 2 (pair0 + q) + 1 (0 past ntz).
template <class Load>
GPVI_SPEC_FN void rows_repack(Cx* X, int np, int ntx, int ntz, int pair0, Load sp, int tid, int nthr)
{
    chunked(
        np * ntx, tid, nthr,
        [&](int t) {
            const int q = t / ntx, f = t - q * ntx;
            const int iz = 2 * (pair0 + q);
            const bool hi = 2 * f > ntx;
            const int g = hi ? ntx - f : f;
            Cx a = iz < ntz ? sp(iz, g) : cx(0.0, 0.0);
            Cx b = iz + 1 < ntz ? sp(iz + 1, g) : cx(0.0, 0.0);
            if (g == 0 || 2 * g == ntx) {
                a.im = 0.0;
                b.im = 0.0;
            }
            if (hi) {
                a.im = -a.im;
                b.im = -b.im;
            }
            return cx(a.re - b.im, a.im + b.re);
        },
        [&](int t, Cx v) { X[t] = v; });
}

}  // namespace spec
}  // namespace gpvi
