// Complex one-dimensional Hermite coefficients of a primitive pair.
#pragma once

#include "gh_common.h"

namespace gh {

// Binomial coefficient C(n, k) (exact in double for the orders used here).
__host__ __device__ constexpr double binom(int n, int k) {
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (double)(n - k + i) / (double)i;
    return r;
}

// Complex 1-D Hermite table Et^{ij}[k] of one primitive pair on one axis:
// E^{ij}_t from the field-free recurrence (hermite.cpp:8-40: raise the bra
// power i times with P-A, then the ket power j times with P-B), then folded
// with the pair's London wave number q: Et[k] = sum_t E_t C(t,k) (-iq)^(t-k).
//
// Pairs with i + j < NR run in registers (fully unrolled over NR entries);
// larger pairs work in place in dst[0 .. i+j] (the real E_t in dst[t].x
// during the recurrence, the fold overwriting dst[k] once E_k is no longer
// needed), so no per-thread array is sized for the largest shell.
constexpr int kHermiteRegs = 7;  // i + j <= 6: shell pairs up to (f, f)

template <int NR>
__device__ inline void phase_hermite_regs(int i, int j, double PA, double PB, double p, double q, zc* dst) {
    double cur[NR];
#pragma unroll
    for (int t = 0; t < NR; ++t) cur[t] = 0.0;
    cur[0] = 1.0;
    const double i2p = 0.5 / p;
    int active = 0;
    for (int step = 0; step < i + j; ++step) {
        const double disp = step < i ? PA : PB;
        double nxt[NR];
#pragma unroll
        for (int t = 0; t < NR; ++t) {
            const double lower = t == 0 ? 0.0 : cur[t - 1];
            const double same = t <= active ? cur[t] : 0.0;
            const double upper = (t < active && t + 1 < NR) ? cur[t + 1 < NR ? t + 1 : NR - 1] : 0.0;
            nxt[t] = t <= active + 1 ? i2p * lower + disp * same + (double)(t + 1) * upper : 0.0;
        }
#pragma unroll
        for (int t = 0; t < NR; ++t) cur[t] = nxt[t];
        ++active;
    }
    const int top = i + j;
    double qpow[NR];
    qpow[0] = 1.0;
#pragma unroll
    for (int m = 1; m < NR; ++m) qpow[m] = qpow[m - 1] * q;
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int m = 0; k + m < NR; ++m) {
            const double v = (k + m <= top) ? cur[k + m] * binom(k + m, k) * qpow[m] : 0.0;
            if ((m & 3) == 0) re += v;
            else if ((m & 3) == 1) im -= v;
            else if ((m & 3) == 2) re -= v;
            else im += v;
        }
        if (k <= top) dst[k] = zc{re, im};
    }
}

__device__ inline void phase_hermite_inplace(int i, int j, double PA, double PB, double p, double q, zc* dst) {
    const int top = i + j;
    const double i2p = 0.5 / p;
    dst[0].x = 1.0;
    for (int step = 0; step < top; ++step) {
        const double disp = step < i ? PA : PB;
        // E_t <- E_{t-1}/(2p) + disp E_t + (t+1) E_{t+1}, from the top down;
        // `upper` is the old E_{t+1}
        double upper = 0.0;
        for (int t = step + 1; t >= 0; --t) {
            const double same = t <= step ? dst[t].x : 0.0;
            const double lower = t > 0 ? dst[t - 1].x : 0.0;
            dst[t].x = i2p * lower + disp * same + (double)(t + 1) * upper;
            upper = same;
        }
    }
    // Et[k] = sum_m E_{k+m} C(k+m, k) (-i q)^m, increasing k
    for (int k = 0; k <= top; ++k) {
        double re = 0.0, im = 0.0, c = 1.0, qp = 1.0;
        for (int m = 0; k + m <= top; ++m) {
            if (m > 0) {
                c = c * (double)(k + m) / (double)m;  // C(k+m, k), exact integers
                qp *= q;
            }
            const double v = dst[k + m].x * c * qp;
            if ((m & 3) == 0) re += v;
            else if ((m & 3) == 1) im -= v;
            else if ((m & 3) == 2) re -= v;
            else im += v;
        }
        dst[k] = zc{re, im};
    }
}

__device__ inline void build_phase_hermite(int i, int j, double PA, double PB, double p, double q, zc* dst) {
    if (i + j < kHermiteRegs) phase_hermite_regs<kHermiteRegs>(i, j, PA, PB, p, q, dst);
    else phase_hermite_inplace(i, j, PA, PB, p, q, dst);
}

}  // namespace gh
