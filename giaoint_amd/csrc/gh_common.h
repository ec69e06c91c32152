// Shared host/device definitions of the gfx950 ERI library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gh {

// Complex double in (re, im) layout, identical to std::complex<double>.
struct zc {
    double x, y;
};

__host__ __device__ inline zc zmake(double x, double y) { return zc{x, y}; }
__host__ __device__ inline zc zadd(zc a, zc b) { return zc{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline zc zsub(zc a, zc b) { return zc{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline zc zscale(double s, zc a) { return zc{s * a.x, s * a.y}; }
__host__ __device__ inline zc zmul(zc a, zc b) { return zc{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ inline zc zneg(zc a) { return zc{-a.x, -a.y}; }
// acc += a * b with fused multiply-adds.
__device__ inline void zmac(zc& acc, zc a, zc b) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
}
__device__ inline zc zexp(zc z) {
    double s, c;
    sincos(z.y, &s, &c);
    const double e = exp(z.x);
    return zc{e * c, e * s};
}
__device__ inline double zabs(zc z) { return hypot(z.x, z.y); }
// Smith's algorithm for a / b.
__device__ inline zc zdiv(zc a, zc b) {
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return zc{(a.x + a.y * r) / d, (a.y - a.x * r) / d};
    }
    const double r = b.x / b.y, d = b.y + b.x * r;
    return zc{(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}
// Principal square root.
__device__ inline zc zsqrt(zc z) {
    if (z.x == 0.0 && z.y == 0.0) return zc{0.0, z.y};
    const double r = hypot(z.x, z.y);
    if (z.x >= 0.0) {
        const double t = sqrt(0.5 * (r + z.x));
        return zc{t, z.y / (2.0 * t)};
    }
    const double t = sqrt(0.5 * (r - z.x));
    return zc{fabs(z.y) / (2.0 * t), copysign(t, z.y)};
}

constexpr double kPi = 3.141592653589793;
constexpr int kMaxShellL = 8;      // per-shell angular momentum handled by the kernels
constexpr int kMaxTotalL = 24;     // combined angular order handled by the kernels
constexpr int kMaxBoysOrder = 32;  // boys.cpp:16

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nherm(int l) { return (l + 1) * (l + 2) * (l + 3) / 6; }

// Per-shell device record (basis upload).
struct DevShell {
    double center[3];
    int l, nprim, ncontr, ncart;
    int prim_off;  // into exps
    int coef_off;  // into coefs: [contraction][primitive]
    int norm_off;  // into norms: [component][primitive]
    int ao_count;
};

// One primitive-quartet batch entry of the seed kernel: a shell quartet.
struct SeedQuartet {
    int sh[4];
    int L;
    int kab, kcd;      // primitive pair counts Ka*Kb, Kc*Kd
    int pad;
    long long pq_begin;  // first flat primitive-quartet index
    long long seed_off;  // offset (in zc) of this quartet's seeds
};

// One unit of the transform kernel.
struct WorkItem {
    int sh[4];
    int x[4];        // contraction rows
    int cd0, cd1;    // ket component range [cd0, cd1)
    long long seed_off;
    long long out_base;     // complex element offset of component (0,0,0,0)
    long long stride[4];    // complex element strides of the four AO indices
    int cls;                // 0: the launch's class; else la | lb<<4 | lc<<8 | ld<<12 | tile<<16
    int pad;
};

// Device error record: first failing lane wins.
struct DevError {
    int code;
    int order;
    long long item;
    double zr, zi;
};

enum DevErrorCode {
    kErrNone = 0,
    kErrBoysDomain = 1,
    kErrBoysSeries = 2,
    kErrBoysQuadrature = 3,
    kErrBoysScaling = 4,
    kErrBoysUnscaled = 5,
};

}  // namespace gh
