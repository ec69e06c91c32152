// Complex one-dimensional Hermite coefficients of a primitive pair.
#pragma once

#include "gh_common.h"

namespace gh {

constexpr int kMaxE = 2 * kMaxShellL + 1;

// Binomial coefficient C(n, k) (exact in double for n <= 2 * kMaxShellL).
__host__ __device__ constexpr double binom(int n, int k) {
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (double)(n - k + i) / (double)i;
    return r;
}

// Complex 1-D Hermite table Et^{ij}[k] of one primitive pair on one axis:
// E^{ij}_t from the field-free recurrence (hermite.cpp:8-40: raise the bra
// power i times with P-A, then the ket power j times with P-B), then folded
// with the pair's London wave number q: Et[k] = sum_t E_t C(t,k) (-iq)^(t-k).
__device__ inline void build_phase_hermite(int i, int j, double PA, double PB, double p, double q, zc* dst) {
    double cur[kMaxE + 1];
#pragma unroll
    for (int t = 0; t <= kMaxE; ++t) cur[t] = 0.0;
    cur[0] = 1.0;
    const double i2p = 0.5 / p;
    int active = 0;
    for (int step = 0; step < i + j; ++step) {
        const double disp = step < i ? PA : PB;
        double nxt[kMaxE + 1];
#pragma unroll
        for (int t = 0; t <= kMaxE; ++t) {
            const double lower = t == 0 ? 0.0 : cur[t - 1];
            const double same = t <= active ? cur[t] : 0.0;
            const double upper = (t < active && t < kMaxE) ? cur[t < kMaxE ? t + 1 : kMaxE] : 0.0;
            nxt[t] = t <= active + 1 ? i2p * lower + disp * same + (double)(t + 1) * upper : 0.0;
        }
#pragma unroll
        for (int t = 0; t <= kMaxE; ++t) cur[t] = nxt[t];
        ++active;
    }
    // powers of (-i q): (-i)^m q^m
    const int top = i + j;
    double qpow[kMaxE];
    qpow[0] = 1.0;
#pragma unroll
    for (int m = 1; m < kMaxE; ++m) qpow[m] = qpow[m - 1] * q;
#pragma unroll
    for (int k = 0; k < kMaxE; ++k) {
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int m = 0; k + m < kMaxE; ++m) {
            const double v = (k + m <= top) ? cur[k + m] * binom(k + m, k) * qpow[m] : 0.0;
            if ((m & 3) == 0) re += v;
            else if ((m & 3) == 1) im -= v;
            else if ((m & 3) == 2) re -= v;
            else im += v;
        }
        if (k <= top) dst[k] = zc{re, im};
    }
}

}  // namespace gh
