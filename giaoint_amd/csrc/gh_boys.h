// Complex Boys function F_n(z) = int_0^1 t^{2n} exp(-z t^2) dt on gfx950.
//
// Same region scheme and tolerances as the reference (boys.cpp:16-20,245-294):
//   |z| <= 0.75 ..................... compensated power series per order
//   Re z >= 2*nmax+60, |Im| <= Re/2 .. positive-sector asymptotic
//   |z| > 160 otherwise ............. BoysNumericalError
//   Re z < 0 ........................ scaled e^z F_n by adaptive quadrature
//   otherwise ....................... adaptive 16/32-point Gauss-Legendre
// The adaptive quadrature starts at ceil(max(|Re|,|Im|)/8) segments and
// doubles up to 4096 until the two rules agree within 2e-15 + 7.5e-14 |F|.
// One lane evaluates one argument; orders live in registers (unrolled to
// NMAX, predicated on the runtime nmax <= NMAX).
#pragma once

#include "gh_common.h"

namespace gh {

extern __constant__ double c_gl16x[16];
extern __constant__ double c_gl16w[16];
extern __constant__ double c_gl32x[32];
extern __constant__ double c_gl32w[32];

constexpr double kBoysAtol = 2.0e-15;
constexpr double kBoysRtol = 7.5e-14;

// Compensated power series of every order n <= nmax. Each order sums
// (-z)^k / k! / (2n + 2k + 1); the powers (-z)^k / k! are formed once for all
// orders by the recurrence each order would use on its own, so every order's
// terms, sums and stopping index are those of its own series.
template <int NMAX>
__device__ inline int boys_series(zc z, int nmax, bool scaled, zc (&F)[NMAX + 1], double& worst) {
    worst = 0.0;
    const zc mz = zneg(z);
    zc sum[NMAX + 1], comp[NMAX + 1];
    double last[NMAX + 1];
    unsigned long long open = 0;  // bit n: order n still summing
#pragma unroll
    for (int n = 0; n <= NMAX; ++n) {
        sum[n] = comp[n] = zc{0.0, 0.0};
        last[n] = 0.0;
        if (n <= nmax) open |= 1ull << n;
    }
    zc pof{1.0, 0.0};
    for (int k = 0; k < 256 && open; ++k) {
#pragma unroll
        for (int n = 0; n <= NMAX; ++n) {
            if (!((open >> n) & 1ull)) continue;
            const double den = (double)(2 * n + 2 * k + 1);
            const zc term{pof.x / den, pof.y / den};
            const zc corrected = zsub(term, comp[n]);
            const zc next = zadd(sum[n], corrected);
            comp[n] = zsub(zsub(next, sum[n]), corrected);
            sum[n] = next;
            last[n] = zabs(term);
            if (k > 2 && last[n] <= kBoysAtol + kBoysRtol * zabs(sum[n])) open &= ~(1ull << n);
        }
        const double kk = (double)(k + 1);
        pof = zmul(pof, zc{mz.x / kk, mz.y / kk});
    }
    const zc ez = scaled ? zexp(z) : zc{1.0, 0.0};
#pragma unroll
    for (int n = 0; n <= NMAX; ++n) {
        if (n > nmax) continue;
        F[n] = scaled ? zmul(ez, sum[n]) : sum[n];
        worst = fmax(worst, last[n]);
    }
    return open == 0 ? kErrNone : kErrBoysSeries;
}

// One pass of composite 16- and 32-point Gauss-Legendre over `segments`.
template <int NMAX>
__device__ inline void boys_quadrature_pass(zc z, int nmax, bool scaled, int segments, zc (&lo)[NMAX + 1],
                                            zc (&hi)[NMAX + 1]) {
#pragma unroll
    for (int n = 0; n <= NMAX; ++n) lo[n] = hi[n] = zc{0.0, 0.0};
    const double inv = 1.0 / (double)segments;
    const double hw = 0.5 * inv;
    for (int s = 0; s < segments; ++s) {
        const double mid = ((double)s + 0.5) * inv;
        for (int p = 0; p < 16; ++p) {
            const double t = mid + hw * c_gl16x[p];
            const double t2 = t * t;
            const zc e = scaled ? zscale(1.0 - t2, z) : zscale(t2, zneg(z));
            const zc c = zscale(hw * c_gl16w[p], zexp(e));
            double pw = 1.0;
#pragma unroll
            for (int n = 0; n <= NMAX; ++n) {
                if (n <= nmax) {
                    lo[n].x += c.x * pw;
                    lo[n].y += c.y * pw;
                }
                pw *= t2;
            }
        }
        for (int p = 0; p < 32; ++p) {
            const double t = mid + hw * c_gl32x[p];
            const double t2 = t * t;
            const zc e = scaled ? zscale(1.0 - t2, z) : zscale(t2, zneg(z));
            const zc c = zscale(hw * c_gl32w[p], zexp(e));
            double pw = 1.0;
#pragma unroll
            for (int n = 0; n <= NMAX; ++n) {
                if (n <= nmax) {
                    hi[n].x += c.x * pw;
                    hi[n].y += c.y * pw;
                }
                pw *= t2;
            }
        }
    }
}

template <int NMAX>
__device__ inline int boys_adaptive(zc z, int nmax, bool scaled, zc (&F)[NMAX + 1], double& err, int& segs) {
    const double difficulty = fmax(fabs(z.x), fabs(z.y));
    int segments = (int)ceil(difficulty / 8.0);
    if (segments < 1) segments = 1;
    zc lo[NMAX + 1];
    for (;;) {
        boys_quadrature_pass<NMAX>(z, nmax, scaled, segments, lo, F);
        bool converged = true;
        double worst = 0.0;
#pragma unroll
        for (int n = 0; n <= NMAX; ++n) {
            if (n <= nmax) {
                const double e = zabs(zsub(F[n], lo[n]));
                worst = fmax(worst, e);
                converged = converged && e <= kBoysAtol + kBoysRtol * zabs(F[n]);
            }
        }
        if (converged) {
            err = worst;
            segs = segments;
            return kErrNone;
        }
        if (segments >= 4096) return kErrBoysQuadrature;
        segments *= 2;
    }
}

template <int NMAX>
__device__ inline int boys_asymptotic(zc z, int nmax, bool scaled, zc (&F)[NMAX + 1], double& err) {
    zc inv = zdiv(zc{1.0, 0.0}, zsqrt(z));
    double gamma = sqrt(kPi);
    const zc ez = scaled ? zexp(z) : zc{1.0, 0.0};
    bool finite = true;
#pragma unroll
    for (int n = 0; n <= NMAX; ++n) {
        if (n > nmax) continue;
        zc v = zscale(0.5 * gamma, inv);
        if (scaled) v = zmul(v, ez);
        finite = finite && isfinite(v.x) && isfinite(v.y);
        F[n] = v;
        gamma *= (double)n + 0.5;
        inv = zdiv(inv, z);
    }
    if (!finite) return kErrBoysScaling;
    err = zabs(zexp(zneg(z))) / fmax(1.0, 2.0 * zabs(z));
    return kErrNone;
}

// Returns a DevErrorCode; region follows giao::BoysRegion numbering.
template <int NMAX>
__device__ inline int boys_eval(zc z, int nmax, bool scaled, zc (&F)[NMAX + 1], int& region, double& err,
                                int& segs) {
    segs = 0;
    err = 0.0;
    const double mag = zabs(z);
    if (mag <= 0.75) {
        region = 0;
        return boys_series<NMAX>(z, nmax, scaled, F, err);
    }
    const double asym = 2.0 * (double)nmax + 60.0;
    if (z.x >= asym && fabs(z.y) <= 0.5 * z.x) {
        region = 3;
        return boys_asymptotic<NMAX>(z, nmax, scaled, F, err);
    }
    if (mag > 160.0) return kErrBoysDomain;
    if (z.x < 0.0) {
        region = 2;
        const int st = boys_adaptive<NMAX>(z, nmax, true, F, err, segs);
        if (st != kErrNone || scaled) return st;
        const zc factor = zexp(zneg(z));
#pragma unroll
        for (int n = 0; n <= NMAX; ++n) {
            if (n <= nmax) {
                F[n] = zmul(F[n], factor);
                if (!isfinite(F[n].x) || !isfinite(F[n].y)) return kErrBoysUnscaled;
            }
        }
        err *= zabs(factor);
        return kErrNone;
    }
    region = 1;
    return boys_adaptive<NMAX>(z, nmax, scaled, F, err, segs);
}

}  // namespace gh
