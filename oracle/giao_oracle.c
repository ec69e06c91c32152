/*
 * CPU oracle: plain-C restatement of GIAOINT's London-orbital ERI path.
 *
 * TEST INFRASTRUCTURE ONLY (see giao_oracle.h). Every function cites the
 * reference file:line it follows; arithmetic is written in the reference's
 * operation order so that results agree with the compiled reference to a few
 * ulps. The per-component structure (one six-index complex R table per
 * Cartesian component quartet, Boys cached per argument) is deliberately the
 * reference's, so that this file also serves as the timed CPU baseline.
 */
#include "giao_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define PI 3.141592653589793
#define BOYS_MAX_ORDER 32          /* boys.cpp:16 */
#define BOYS_DIRECT_LIMIT 160.0    /* boys.cpp:17 */
#define BOYS_SERIES_LIMIT 0.75     /* boys.cpp:18 */
#define BOYS_ATOL 2.0e-15          /* boys.cpp:19 */
#define BOYS_RTOL 7.5e-14          /* boys.cpp:20 */
#define MAX_AUX_ENTRIES 4000000u   /* eri.cpp:24 */

static __thread char g_msg[512];

const char* or_last_error(void) { return g_msg; }

static int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof g_msg, fmt, ap);
    va_end(ap);
    return code;
}

/* ------------------------------------------------------------------------ */
/* types.cpp                                                                */
/* ------------------------------------------------------------------------ */

/* types.cpp:168-175 */
static double log_odd_double_factorial(int l) {
    if (l == 0) return 0.0;
    const double n = (double)l;
    return n * log(2.0) + lgamma(n + 0.5) - 0.5 * log(PI);
}

/* types.cpp:207-219 */
double or_primitive_normalization(double exponent, int ax, int ay, int az) {
    const double total = (double)(ax + ay + az);
    const double log_den = log_odd_double_factorial(ax) +
                           log_odd_double_factorial(ay) +
                           log_odd_double_factorial(az);
    const double log_norm = 0.75 * log(2.0 * exponent / PI) +
                            0.5 * (total * log(4.0 * exponent) - log_den);
    return exp(log_norm);
}

/* types.cpp:177-188 */
static double same_center_overlap(double alpha, double beta, int l) {
    const double norm = or_primitive_normalization(alpha, l, 0, 0) *
                        or_primitive_normalization(beta, l, 0, 0);
    const double p = alpha + beta;
    const double lm = lgamma((double)l + 0.5) - ((double)l + 0.5) * log(p);
    return norm * exp(lm) * PI / p;
}

/* types.cpp:239-306 (ContractionNormalization::normalize) */
int or_normalize_shell(int L, int nprim, int ncontr, const double* exps,
                       const double* coefs_in, double* coefs_out) {
    if (nprim < 1) return fail(OR_INVALID_ARGUMENT, "a shell requires at least one primitive");
    if (ncontr < 1) return fail(OR_INVALID_ARGUMENT, "a shell requires at least one contraction");
    for (int i = 0; i < nprim; ++i)
        if (!isfinite(exps[i]) || exps[i] <= 0.0)
            return fail(OR_INVALID_ARGUMENT, "shell exponents must be finite and positive");
    for (int i = 0; i < nprim * ncontr; ++i)
        if (!isfinite(coefs_in[i]))
            return fail(OR_INVALID_ARGUMENT, "shell coefficients must be finite");
    if (coefs_out != coefs_in) memcpy(coefs_out, coefs_in, sizeof(double) * (size_t)(nprim * ncontr));
    for (int x = 0; x < ncontr; ++x) {
        const double* c = coefs_out + (size_t)x * nprim;
        int any = 0;
        for (int i = 0; i < nprim; ++i) any = any || c[i] != 0.0;
        if (!any) return fail(OR_INVALID_ARGUMENT, "a shell contraction cannot have all-zero coefficients");
        double norm2 = 0.0;
        for (int i = 0; i < nprim; ++i)
            for (int j = 0; j < nprim; ++j)
                norm2 += c[i] * c[j] * same_center_overlap(exps[i], exps[j], L);
        if (!isfinite(norm2) || norm2 <= 0.0)
            return fail(OR_INVALID_ARGUMENT, "contraction has a non-positive or non-finite norm");
        const double scale = 1.0 / sqrt(norm2);
        for (int i = 0; i < nprim; ++i) coefs_out[(size_t)x * nprim + i] *= scale;
    }
    return OR_OK;
}

/* types.cpp:221-237: x-major Cartesian order. */
static int cartesian_components(int l, int (*out)[3]) {
    int k = 0;
    for (int x = l; x >= 0; --x) {
        const int rem = l - x;
        for (int y = rem; y >= 0; --y) {
            out[k][0] = x;
            out[k][1] = y;
            out[k][2] = rem - y;
            ++k;
        }
    }
    return k;
}

/* types.cpp:326-333 */
static void london_wave_vector(const double field[6], const double c[3], double k[3]) {
    const double dx = c[0] - field[3], dy = c[1] - field[4], dz = c[2] - field[5];
    const double Bx = field[0], By = field[1], Bz = field[2];
    k[0] = 0.5 * (By * dz - Bz * dy);
    k[1] = 0.5 * (Bz * dx - Bx * dz);
    k[2] = 0.5 * (Bx * dy - By * dx);
}

/* ------------------------------------------------------------------------ */
/* overlap.cpp:60-91 gaussian_product                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    double exponent, reduced;
    double P[3];
    double q[3];
    cplx Pc[3];
    double ordinary;
    cplx london;
} pair_data;

static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void gaussian_product(double ea, const double* A, double eb, const double* B,
                             const double field[6], pair_data* out) {
    const double p = ea + eb;
    out->exponent = p;
    out->reduced = ea * eb / p;
    for (int i = 0; i < 3; ++i) out->P[i] = (ea * A[i] + eb * B[i]) / p;
    double ka[3], kb[3];
    london_wave_vector(field, A, ka);
    london_wave_vector(field, B, kb);
    for (int i = 0; i < 3; ++i) out->q[i] = kb[i] - ka[i];
    const double d[3] = {A[0] - B[0], A[1] - B[1], A[2] - B[2]};
    out->ordinary = exp(-out->reduced * dot3(d, d));
    const double phase = -dot3(out->q, out->P);
    const double damping = -dot3(out->q, out->q) / (4.0 * p);
    out->london = cexp(CMPLX(damping, phase));
    for (int i = 0; i < 3; ++i) out->Pc[i] = CMPLX(out->P[i], -0.5 * out->q[i] / p);
}

/* ------------------------------------------------------------------------ */
/* boys.cpp                                                                 */
/* ------------------------------------------------------------------------ */

typedef struct { double nodes[32]; double weights[32]; int n; } gl_rule;
static gl_rule g_rule16, g_rule32;

/* boys.cpp:27-109: Newton iteration on Legendre polynomials. */
static void build_rule(gl_rule* r, int n) {
    r->n = n;
    for (int i = 0; i < n / 2; ++i) {
        double root = cos(PI * ((double)i + 0.75) / ((double)n + 0.5));
        double derivative = 0.0;
        for (int it = 0; it < 20; ++it) {
            double prev = 1.0, cur = root;
            for (int deg = 2; deg <= n; ++deg) {
                const double next = ((2.0 * (double)deg - 1.0) * root * cur - ((double)deg - 1.0) * prev) / (double)deg;
                prev = cur;
                cur = next;
            }
            derivative = (double)n * (root * cur - prev) / (root * root - 1.0);
            const double update = cur / derivative;
            root -= update;
            if (fabs(update) < 4.0 * 2.220446049250313e-16) break;
        }
        const double w = 2.0 / ((1.0 - root * root) * derivative * derivative);
        r->nodes[i] = -root;
        r->nodes[n - 1 - i] = root;
        r->weights[i] = w;
        r->weights[n - 1 - i] = w;
    }
}

__attribute__((constructor)) static void init_rules(void) {
    build_rule(&g_rule16, 16);
    build_rule(&g_rule32, 32);
}

/* boys.cpp:111-135 */
static void quadrature(cplx z, int scaled, size_t segments, const gl_rule* r, cplx* v, int nv) {
    for (int i = 0; i < nv; ++i) v[i] = 0.0;
    const double inv = 1.0 / (double)segments;
    for (size_t s = 0; s < segments; ++s) {
        const double mid = ((double)s + 0.5) * inv;
        const double hw = 0.5 * inv;
        for (int p = 0; p < r->n; ++p) {
            const double t = mid + hw * r->nodes[p];
            const double t2 = t * t;
            const cplx e = scaled ? z * (1.0 - t2) : -z * t2;
            const cplx common = hw * r->weights[p] * cexp(e);
            double power = 1.0;
            for (int i = 0; i < nv; ++i) {
                v[i] += common * power;
                power *= t2;
            }
        }
    }
}

/* boys.cpp:137-169 */
static int power_series(cplx z, int scaled, cplx* v, int nv, double* worst_out) {
    double worst = 0.0;
    for (int n = 0; n < nv; ++n) {
        cplx pof = 1.0, sum = 0.0, comp = 0.0;
        double last = 0.0;
        for (int k = 0; k < 256; ++k) {
            const cplx term = pof / (double)(2 * n + 2 * k + 1);
            const cplx corrected = term - comp;
            const cplx next = sum + corrected;
            comp = (next - sum) - corrected;
            sum = next;
            last = cabs(term);
            if (k > 2 && last <= BOYS_ATOL + BOYS_RTOL * cabs(sum)) break;
            pof *= -z / (double)(k + 1);
            if (k == 255) return fail(OR_BOYS_ERROR, "complex Boys power series did not converge");
        }
        v[n] = scaled ? cexp(z) * sum : sum;
        if (last > worst) worst = last;
    }
    *worst_out = worst;
    return OR_OK;
}

/* boys.cpp:171-207 */
static int adaptive_quadrature(cplx z, int scaled, cplx* v, int nv, double* err, size_t* seg_out) {
    const double difficulty = fmax(fabs(creal(z)), fabs(cimag(z)));
    size_t segments = (size_t)ceil(difficulty / 8.0);
    if (segments < 1) segments = 1;
    cplx low[BOYS_MAX_ORDER + 1], high[BOYS_MAX_ORDER + 1];
    for (;;) {
        quadrature(z, scaled, segments, &g_rule16, low, nv);
        quadrature(z, scaled, segments, &g_rule32, high, nv);
        int converged = 1;
        double worst = 0.0;
        for (int n = 0; n < nv; ++n) {
            const double e = cabs(high[n] - low[n]);
            if (e > worst) worst = e;
            converged = converged && e <= BOYS_ATOL + BOYS_RTOL * cabs(high[n]);
        }
        if (converged) {
            for (int n = 0; n < nv; ++n) v[n] = high[n];
            *err = worst;
            *seg_out = segments;
            return OR_OK;
        }
        if (segments >= 4096u) return fail(OR_BOYS_ERROR, "adaptive complex Boys quadrature did not reach tolerance");
        segments *= 2u;
    }
}

/* boys.cpp:209-231 */
static int positive_asymptotic(cplx z, int scaled, cplx* v, int nv, double* err) {
    const cplx root = csqrt(z);
    cplx inv_pow = 1.0 / root;
    double gamma = sqrt(PI);
    for (int n = 0; n < nv; ++n) {
        cplx value = 0.5 * gamma * inv_pow;
        if (scaled) value *= cexp(z);
        if (!isfinite(creal(value)) || !isfinite(cimag(value)))
            return fail(OR_BOYS_ERROR, "requested Boys scaling is not representable for this argument");
        v[n] = value;
        gamma *= (double)n + 0.5;
        inv_pow /= z;
    }
    *err = cabs(cexp(-z)) / fmax(1.0, 2.0 * cabs(z));
    return OR_OK;
}

/* boys.cpp:245-294 */
static int compute_boys(cplx z, int nmax, int scaled, cplx* v, int* region, double* err, size_t* segs) {
    if (nmax < 0) return fail(OR_INVALID_ARGUMENT, "Boys output must contain at least F_0");
    if (nmax > BOYS_MAX_ORDER) return fail(OR_INVALID_ARGUMENT, "complex Boys supports orders through 32");
    if (!isfinite(creal(z)) || !isfinite(cimag(z))) return fail(OR_INVALID_ARGUMENT, "complex Boys argument must be finite");
    const int nv = nmax + 1;
    *segs = 0;
    const double mag = cabs(z);
    if (mag <= BOYS_SERIES_LIMIT) {
        *region = 0;
        return power_series(z, scaled, v, nv, err);
    }
    const double asym = 2.0 * (double)nmax + 60.0;
    if (creal(z) >= asym && fabs(cimag(z)) <= 0.5 * creal(z)) {
        *region = 3;
        return positive_asymptotic(z, scaled, v, nv, err);
    }
    if (mag > BOYS_DIRECT_LIMIT)
        return fail(OR_BOYS_ERROR,
                    "complex Boys argument (%.17g,%.17g) at order %d is outside the verified direct domain "
                    "|z| <= 160 and the conservative large-positive asymptotic sector",
                    creal(z), cimag(z), nmax);
    if (creal(z) < 0.0) {
        *region = 2;
        if (scaled) return adaptive_quadrature(z, 1, v, nv, err, segs);
        int st = adaptive_quadrature(z, 1, v, nv, err, segs);
        if (st) return st;
        const cplx factor = cexp(-z);
        for (int n = 0; n < nv; ++n) {
            v[n] *= factor;
            if (!isfinite(creal(v[n])) || !isfinite(cimag(v[n])))
                return fail(OR_BOYS_ERROR, "unscaled Boys value is not representable for this argument");
        }
        *err *= cabs(factor);
        return OR_OK;
    }
    *region = 1;
    return adaptive_quadrature(z, scaled, v, nv, err, segs);
}

int or_boys(double zr, double zi, int nmax, int scaled, double* out, int* region,
            double* error_estimate, int* segments) {
    cplx v[BOYS_MAX_ORDER + 1];
    int reg = -1;
    double err = 0.0;
    size_t segs = 0;
    const int st = compute_boys(CMPLX(zr, zi), nmax, scaled, v, &reg, &err, &segs);
    if (st) return st;
    for (int n = 0; n <= nmax; ++n) {
        out[2 * n] = creal(v[n]);
        out[2 * n + 1] = cimag(v[n]);
    }
    if (region) *region = reg;
    if (error_estimate) *error_estimate = err;
    if (segments) *segments = (int)segs;
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* hermite.cpp:8-40                                                         */
/* ------------------------------------------------------------------------ */

static void hermite_coefficients(int la, int lb, double A, double B, double P, double p, double* out) {
    const int maxo = la + lb;
    double cur[2 * BOYS_MAX_ORDER + 4], nxt[2 * BOYS_MAX_ORDER + 4];
    for (int t = 0; t < maxo + 2; ++t) cur[t] = 0.0;
    cur[0] = 1.0;
    int active = 0;
    const double i2p = 0.5 / p;
    for (int step = 0; step < la + lb; ++step) {
        const double disp = step < la ? P - A : P - B;
        for (int t = 0; t < maxo + 2; ++t) nxt[t] = 0.0;
        for (int t = 0; t <= active + 1; ++t) {
            const double lower = t == 0 ? 0.0 : cur[t - 1];
            const double same = t <= active ? cur[t] : 0.0;
            const double upper = t < active ? cur[t + 1] : 0.0;
            nxt[t] = i2p * lower + disp * same + (double)(t + 1) * upper;
        }
        ++active;
        memcpy(cur, nxt, sizeof(double) * (size_t)(maxo + 2));
    }
    for (int t = 0; t <= maxo; ++t) out[t] = cur[t];
}

/* ------------------------------------------------------------------------ */
/* eri.cpp:67-384 EriKernel::compute                                        */
/* ------------------------------------------------------------------------ */

typedef struct {
    double exponent;
    double center[3];
    int ang[3];
    double coefficient;
    int normalized;
} prim_t;

typedef struct {
    /* Boys cache (eri.cpp:173-187) */
    int boys_valid;
    cplx boys_arg;
    int boys_scaled;
    int boys_order;
    cplx boys[BOYS_MAX_ORDER + 1];
    /* Hermite product cache (eri.cpp:105-160) */
    int herm_valid[2];
    int herm_ang[2][6];
    double herm_geo[2][10];
    double herm1d[6][2 * 16 + 1];
    double *prod[2];
    size_t prod_cap[2];
    cplx* aux;
    size_t aux_cap;
    double* raux;
    size_t raux_cap;
} workspace;

static void ws_free(workspace* w) {
    free(w->prod[0]);
    free(w->prod[1]);
    free(w->aux);
    free(w->raux);
}

static void* grow(void* p, size_t* cap, size_t need, size_t elem) {
    if (need <= *cap) return p;
    void* q = realloc(p, need * elem);
    if (!q) { free(p); *cap = 0; return NULL; }
    *cap = need;
    return q;
}

static int eri_kernel(const prim_t* a, const prim_t* b, const prim_t* c, const prim_t* d,
                      const pair_data* ab, const pair_data* cd, int real_zero, workspace* w, cplx* result) {
    const prim_t* prims[4] = {a, b, c, d};
    const pair_data* pairs[2] = {ab, cd};
    size_t maxima[6];
    for (int i = 0; i < 3; ++i) {
        maxima[i] = (size_t)(a->ang[i] + b->ang[i]);
        maxima[3 + i] = (size_t)(c->ang[i] + d->ang[i]);
    }
    const size_t L = (size_t)(a->ang[0] + a->ang[1] + a->ang[2] + b->ang[0] + b->ang[1] + b->ang[2] +
                              c->ang[0] + c->ang[1] + c->ang[2] + d->ang[0] + d->ang[1] + d->ang[2]);
    if (L > 32) return fail(OR_INVALID_ARGUMENT, "electron repulsion supports combined angular order through 32");
    size_t dims = L + 1;
    for (int i = 0; i < 6; ++i) dims *= maxima[i] + 1;
    if (dims > MAX_AUX_ENTRIES) return fail(OR_LENGTH_ERROR, "ERI auxiliary exceeds the workspace limit");

    for (int pi = 0; pi < 2; ++pi) {
        const prim_t* left = prims[2 * pi];
        const prim_t* right = prims[2 * pi + 1];
        const pair_data* pr = pairs[pi];
        int akey[6] = {left->ang[0], left->ang[1], left->ang[2], right->ang[0], right->ang[1], right->ang[2]};
        double gkey[10] = {left->center[0], left->center[1], left->center[2], right->center[0],
                           right->center[1], right->center[2], pr->P[0], pr->P[1], pr->P[2], pr->exponent};
        if (w->herm_valid[pi] && memcmp(w->herm_ang[pi], akey, sizeof akey) == 0 &&
            memcmp(w->herm_geo[pi], gkey, sizeof gkey) == 0)
            continue;
        for (int ax = 0; ax < 3; ++ax)
            hermite_coefficients(left->ang[ax], right->ang[ax], left->center[ax], right->center[ax],
                                 pr->P[ax], pr->exponent, w->herm1d[3 * pi + ax]);
        memcpy(w->herm_ang[pi], akey, sizeof akey);
        memcpy(w->herm_geo[pi], gkey, sizeof gkey);
        w->herm_valid[pi] = 1;
        const size_t xe = maxima[3 * pi] + 1, ye = maxima[3 * pi + 1] + 1, ze = maxima[3 * pi + 2] + 1;
        w->prod[pi] = grow(w->prod[pi], &w->prod_cap[pi], xe * ye * ze, sizeof(double));
        if (!w->prod[pi]) return fail(OR_RUNTIME_ERROR, "out of memory");
        for (size_t x = 0; x < xe; ++x)
            for (size_t y = 0; y < ye; ++y)
                for (size_t z = 0; z < ze; ++z)
                    w->prod[pi][(x * ye + y) * ze + z] =
                        w->herm1d[3 * pi][x] * w->herm1d[3 * pi + 1][y] * w->herm1d[3 * pi + 2][z];
    }

    const double p = ab->exponent, q = cd->exponent;
    const double rho = p * q / (p + q);
    const cplx disp[3] = {ab->Pc[0] - cd->Pc[0], ab->Pc[1] - cd->Pc[1], ab->Pc[2] - cd->Pc[2]};
    const cplx arg = rho * (disp[0] * disp[0] + disp[1] * disp[1] + disp[2] * disp[2]);
    const int scaled = creal(arg) < 0.0;
    if (!w->boys_valid || w->boys_arg != arg || w->boys_scaled != scaled || w->boys_order < (int)L) {
        int region;
        double err;
        size_t segs;
        const int st = compute_boys(arg, (int)L, scaled, w->boys, &region, &err, &segs);
        if (st) return st;
        w->boys_arg = arg;
        w->boys_scaled = scaled;
        w->boys_order = (int)L;
        w->boys_valid = 1;
    }

    cplx common = ab->london * cd->london;
    if (scaled) {
        const double rd[3] = {ab->P[0] - cd->P[0], ab->P[1] - cd->P[1], ab->P[2] - cd->P[2]};
        const double tw[3] = {ab->q[0] + cd->q[0], ab->q[1] + cd->q[1], ab->q[2] + cd->q[2]};
        const double wc[3] = {(p * ab->P[0] + q * cd->P[0]) / (p + q), (p * ab->P[1] + q * cd->P[1]) / (p + q),
                              (p * ab->P[2] + q * cd->P[2]) / (p + q)};
        const double re = -rho * dot3(rd, rd) - dot3(tw, tw) / (4.0 * (p + q));
        common = cexp(CMPLX(re, -dot3(tw, wc)));
    }

    cplx contraction = 0.0;
    if (L == 0) {
        contraction = common * w->boys[0];
    } else if (real_zero) {
        const size_t comb[3] = {maxima[0] + maxima[3], maxima[1] + maxima[4], maxima[2] + maxima[5]};
        const size_t orders = L + 1;
        const size_t strides[3] = {(comb[1] + 1) * (comb[2] + 1) * orders, (comb[2] + 1) * orders, orders};
        const size_t size = (comb[0] + 1) * strides[0];
        w->raux = grow(w->raux, &w->raux_cap, size, sizeof(double));
        if (!w->raux) return fail(OR_RUNTIME_ERROR, "out of memory");
        double* table = w->raux;
        double rf = 1.0;
        for (size_t n = 0; n <= L; ++n) {
            table[n] = rf * creal(w->boys[n]);
            rf *= -2.0 * rho;
        }
        for (size_t x = 0; x <= comb[0]; ++x)
            for (size_t y = 0; y <= comb[1]; ++y)
                for (size_t z = 0; z <= comb[2]; ++z) {
                    const size_t degree = x + y + z;
                    if (degree == 0) continue;
                    const int axis = x > 0 ? 0 : (y > 0 ? 1 : 2);
                    const size_t power = axis == 0 ? x : (axis == 1 ? y : z);
                    const size_t index = x * strides[0] + y * strides[1] + z * strides[2];
                    const size_t lower = index - strides[axis];
                    const double shift = creal(disp[axis]);
                    for (size_t n = 0; n <= L - degree; ++n) {
                        double value = shift * table[lower + n + 1];
                        if (power > 1) value += (double)(power - 1) * table[lower - strides[axis] + n + 1];
                        table[index + n] = value;
                    }
                }
        double rc = 0.0;
        for (size_t t = 0; t <= maxima[0]; ++t)
            for (size_t u = 0; u <= maxima[1]; ++u)
                for (size_t v = 0; v <= maxima[2]; ++v) {
                    const double cab = w->prod[0][(t * (maxima[1] + 1) + u) * (maxima[2] + 1) + v];
                    for (size_t tau = 0; tau <= maxima[3]; ++tau)
                        for (size_t phi = 0; phi <= maxima[4]; ++phi)
                            for (size_t chi = 0; chi <= maxima[5]; ++chi) {
                                const double sign = (tau + phi + chi) % 2 == 0 ? 1.0 : -1.0;
                                rc += cab * w->prod[1][(tau * (maxima[4] + 1) + phi) * (maxima[5] + 1) + chi] *
                                      (sign * table[(t + tau) * strides[0] + (u + phi) * strides[1] +
                                                    (v + chi) * strides[2]]);
                            }
                }
        contraction = rc;
    } else {
        const size_t orders = L + 1;
        size_t strides[6];
        strides[5] = orders;
        for (int ax = 5; ax > 0; --ax) strides[ax - 1] = strides[ax] * (maxima[ax] + 1);
        const size_t entries = strides[0] * (maxima[0] + 1) / orders;
        w->aux = grow(w->aux, &w->aux_cap, entries * orders, sizeof(cplx));
        if (!w->aux) return fail(OR_RUNTIME_ERROR, "out of memory");
        cplx* aux = w->aux;
        double rf = 1.0;
        for (size_t n = 0; n <= L; ++n) {
            aux[n] = common * rf * w->boys[n];
            rf *= -2.0 * rho;
        }
        const double waves[6] = {ab->q[0], ab->q[1], ab->q[2], cd->q[0], cd->q[1], cd->q[2]};
        for (size_t entry = 1; entry < entries; ++entry) {
            size_t powers[6];
            size_t rem = entry, degree = 0;
            for (int ax = 6; ax-- > 0;) {
                powers[ax] = rem % (maxima[ax] + 1);
                rem /= maxima[ax] + 1;
                degree += powers[ax];
            }
            int axis = 0;
            while (powers[axis] == 0) ++axis;
            const size_t index = entry * orders;
            const size_t lower = index - strides[axis];
            const int coord = axis % 3;
            const cplx shift = axis < 3 ? disp[coord] : -disp[coord];
            const cplx wave = CMPLX(0.0, -waves[axis]);
            for (size_t n = 0; n <= L - degree; ++n) {
                cplx value = wave * aux[lower + n];
                value += shift * aux[lower + n + 1];
                if (powers[axis] > 1) value += (double)(powers[axis] - 1) * aux[lower - strides[axis] + n + 1];
                if (axis < 3 && powers[axis + 3] > 0)
                    value -= (double)powers[axis + 3] * aux[lower - strides[axis + 3] + n + 1];
                aux[index + n] = value;
            }
        }
        const size_t aye = maxima[1] + 1, aze = maxima[2] + 1, cye = maxima[4] + 1, cze = maxima[5] + 1;
        for (size_t t = 0; t <= maxima[0]; ++t)
            for (size_t u = 0; u <= maxima[1]; ++u)
                for (size_t v = 0; v <= maxima[2]; ++v) {
                    const double cab = w->prod[0][(t * aye + u) * aze + v];
                    for (size_t tau = 0; tau <= maxima[3]; ++tau)
                        for (size_t phi = 0; phi <= maxima[4]; ++phi)
                            for (size_t chi = 0; chi <= maxima[5]; ++chi)
                                contraction += cab * w->prod[1][(tau * cye + phi) * cze + chi] *
                                               aux[t * strides[0] + u * strides[1] + v * strides[2] +
                                                   tau * strides[3] + phi * strides[4] + chi * strides[5]];
                }
    }

    const double prefactor = 2.0 * pow(PI, 2.5) / (p * q * sqrt(p + q));
    double nrm[4];
    for (int i = 0; i < 4; ++i)
        nrm[i] = prims[i]->normalized
                     ? or_primitive_normalization(prims[i]->exponent, prims[i]->ang[0], prims[i]->ang[1], prims[i]->ang[2])
                     : 1.0;
    *result = prefactor * ab->ordinary * cd->ordinary * a->coefficient * b->coefficient * c->coefficient *
              d->coefficient * nrm[0] * nrm[1] * nrm[2] * nrm[3] * contraction;
    return OR_OK;
}

static int use_real_path(const double field[6]) { return field[0] == 0.0 && field[1] == 0.0 && field[2] == 0.0; }

/* eri.cpp:438-452 */
int or_primitive_eri(const double prim[4][6], const int ang[4][3], const double field[6], double out[2]) {
    prim_t p[4];
    for (int i = 0; i < 4; ++i) {
        p[i].exponent = prim[i][0];
        p[i].center[0] = prim[i][1];
        p[i].center[1] = prim[i][2];
        p[i].center[2] = prim[i][3];
        p[i].coefficient = prim[i][4];
        p[i].normalized = prim[i][5] != 0.0;
        for (int k = 0; k < 3; ++k) p[i].ang[k] = ang[i][k];
        if (!isfinite(p[i].exponent) || p[i].exponent <= 0.0)
            return fail(OR_INVALID_ARGUMENT, "Gaussian exponent must be finite and positive");
    }
    pair_data ab, cd;
    gaussian_product(p[0].exponent, p[0].center, p[1].exponent, p[1].center, field, &ab);
    gaussian_product(p[2].exponent, p[2].center, p[3].exponent, p[3].center, field, &cd);
    workspace w;
    memset(&w, 0, sizeof w);
    cplx r;
    const int st = eri_kernel(&p[0], &p[1], &p[2], &p[3], &ab, &cd, use_real_path(field), &w, &r);
    ws_free(&w);
    if (st) return st;
    out[0] = creal(r);
    out[1] = cimag(r);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Basis helpers                                                            */
/* ------------------------------------------------------------------------ */

static size_t ncart(int l) { return (size_t)(l + 1) * (size_t)(l + 2) / 2; }

static size_t shell_ao_count(const or_basis* b, int s) { return (size_t)b->ncontr[s] * ncart(b->L[s]); }

static size_t shell_prim_offset(const or_basis* b, int s) {
    size_t o = 0;
    for (int i = 0; i < s; ++i) o += (size_t)b->nprim[i];
    return o;
}

static size_t shell_coef_offset(const or_basis* b, int s) {
    size_t o = 0;
    for (int i = 0; i < s; ++i) o += (size_t)b->nprim[i] * (size_t)b->ncontr[i];
    return o;
}

size_t or_quartet_size(const or_basis* b, const uint32_t q[4]) {
    return shell_ao_count(b, (int)q[0]) * shell_ao_count(b, (int)q[1]) * shell_ao_count(b, (int)q[2]) *
           shell_ao_count(b, (int)q[3]);
}

/* Per-shell view used by the shell drivers. */
typedef struct {
    int l, nprim, ncontr, ncart;
    const double* center;
    const double* exps;
    const double* coefs;
    int comps[153][3]; /* up to l = 16 */
    double* norms;     /* [component][primitive] */
} shell_view;

static int make_shell(const or_basis* b, int s, shell_view* v) {
    v->l = b->L[s];
    if (v->l < 0 || v->l > 16) return fail(OR_INVALID_ARGUMENT, "shell angular momentum out of supported range");
    v->nprim = b->nprim[s];
    v->ncontr = b->ncontr[s];
    v->center = b->centers + 3 * (size_t)s;
    v->exps = b->exps + shell_prim_offset(b, s);
    v->coefs = b->coefs + shell_coef_offset(b, s);
    v->ncart = cartesian_components(v->l, v->comps);
    v->norms = malloc(sizeof(double) * (size_t)v->ncart * (size_t)v->nprim);
    if (!v->norms) return fail(OR_RUNTIME_ERROR, "out of memory");
    for (int c = 0; c < v->ncart; ++c)
        for (int p = 0; p < v->nprim; ++p)
            v->norms[(size_t)c * v->nprim + p] =
                or_primitive_normalization(v->exps[p], v->comps[c][0], v->comps[c][1], v->comps[c][2]);
    return OR_OK;
}

static void make_prim(const shell_view* s, int p, int comp, prim_t* out) {
    out->exponent = s->exps[p];
    for (int i = 0; i < 3; ++i) out->center[i] = s->center[i];
    for (int i = 0; i < 3; ++i) out->ang[i] = comp < 0 ? 0 : s->comps[comp][i];
    out->coefficient = 1.0;
    out->normalized = 0;
}

/* eri.cpp:454-585 compute_eri */
static int compute_eri_shell(const shell_view* sh[4], const double field[6], cplx* out, workspace* w) {
    const size_t na = (size_t)sh[0]->ncontr * sh[0]->ncart, nb = (size_t)sh[1]->ncontr * sh[1]->ncart;
    const size_t nc = (size_t)sh[2]->ncontr * sh[2]->ncart, nd = (size_t)sh[3]->ncontr * sh[3]->ncart;
    for (size_t i = 0; i < na * nb * nc * nd; ++i) out[i] = 0.0;
    const int real_zero = use_real_path(field);
    const shell_view *a = sh[0], *b = sh[1], *c = sh[2], *d = sh[3];
    for (int pa = 0; pa < a->nprim; ++pa)
        for (int pb = 0; pb < b->nprim; ++pb) {
            pair_data ab;
            gaussian_product(a->exps[pa], a->center, b->exps[pb], b->center, field, &ab);
            for (int pc = 0; pc < c->nprim; ++pc)
                for (int pd = 0; pd < d->nprim; ++pd) {
                    pair_data cd;
                    gaussian_product(c->exps[pc], c->center, d->exps[pd], d->center, field, &cd);
                    for (int ca = 0; ca < a->ncart; ++ca)
                        for (int cb = 0; cb < b->ncart; ++cb)
                            for (int cc = 0; cc < c->ncart; ++cc)
                                for (int cdc = 0; cdc < d->ncart; ++cdc) {
                                    prim_t P[4];
                                    make_prim(a, pa, ca, &P[0]);
                                    make_prim(b, pb, cb, &P[1]);
                                    make_prim(c, pc, cc, &P[2]);
                                    make_prim(d, pd, cdc, &P[3]);
                                    cplx pv;
                                    const int st = eri_kernel(&P[0], &P[1], &P[2], &P[3], &ab, &cd, real_zero, w, &pv);
                                    if (st) return st;
                                    pv *= a->norms[(size_t)ca * a->nprim + pa];
                                    pv *= b->norms[(size_t)cb * b->nprim + pb];
                                    pv *= c->norms[(size_t)cc * c->nprim + pc];
                                    pv *= d->norms[(size_t)cdc * d->nprim + pd];
                                    for (int xa = 0; xa < a->ncontr; ++xa) {
                                        const size_t ia = (size_t)xa * a->ncart + ca;
                                        for (int xb = 0; xb < b->ncontr; ++xb) {
                                            const size_t ib = (size_t)xb * b->ncart + cb;
                                            for (int xc = 0; xc < c->ncontr; ++xc) {
                                                const size_t ic = (size_t)xc * c->ncart + cc;
                                                for (int xd = 0; xd < d->ncontr; ++xd) {
                                                    const size_t id = (size_t)xd * d->ncart + cdc;
                                                    out[((ia * nb + ib) * nc + ic) * nd + id] +=
                                                        a->coefs[(size_t)xa * a->nprim + pa] *
                                                        b->coefs[(size_t)xb * b->nprim + pb] *
                                                        c->coefs[(size_t)xc * c->nprim + pc] *
                                                        d->coefs[(size_t)xd * d->nprim + pd] * pv;
                                                }
                                            }
                                        }
                                    }
                                }
                }
        }
    return OR_OK;
}

static int make_all_shells(const or_basis* b, shell_view** out) {
    shell_view* v = calloc((size_t)b->nshell, sizeof(shell_view));
    if (!v) return fail(OR_RUNTIME_ERROR, "out of memory");
    for (int s = 0; s < b->nshell; ++s) {
        const int st = make_shell(b, s, &v[s]);
        if (st) {
            for (int i = 0; i < s; ++i) free(v[i].norms);
            free(v);
            return st;
        }
    }
    *out = v;
    return OR_OK;
}

static void free_all_shells(const or_basis* b, shell_view* v) {
    for (int s = 0; s < b->nshell; ++s) free(v[s].norms);
    free(v);
}

/* eri.cpp:644-767: validate, then OpenMP schedule(dynamic,1); first error wins. */
int or_eri_batch(const or_basis* b, const uint32_t* quartets, size_t n, const double field[6],
                 const uint64_t* offsets, double* out, int threads) {
    if (threads < 1) return fail(OR_INVALID_ARGUMENT, "ERI thread count must be positive");
    for (size_t i = 0; i < 4 * n; ++i)
        if (quartets[i] >= (uint32_t)b->nshell) return fail(OR_OUT_OF_RANGE, "ERI shell-quartet index is out of range");
    shell_view* sv;
    int st = make_all_shells(b, &sv);
    if (st) return st;
    int first_error = OR_OK;
    char first_msg[512] = {0};
#pragma omp parallel num_threads(threads)
    {
        workspace w;
        memset(&w, 0, sizeof w);
#pragma omp for schedule(dynamic, 1)
        for (long long i = 0; i < (long long)n; ++i) {
            int failed;
#pragma omp atomic read
            failed = first_error;
            if (failed) continue;
            const uint32_t* q = quartets + 4 * i;
            const shell_view* sh[4] = {&sv[q[0]], &sv[q[1]], &sv[q[2]], &sv[q[3]]};
            const int s = compute_eri_shell(sh, field, (cplx*)(out + 2 * offsets[i]), &w);
            if (s) {
#pragma omp critical
                {
                    if (!first_error) {
                        first_error = s;
                        memcpy(first_msg, g_msg, sizeof first_msg);
                    }
                }
            }
        }
        ws_free(&w);
    }
    free_all_shells(b, sv);
    if (first_error) {
        memcpy(g_msg, first_msg, sizeof g_msg);
        return first_error;
    }
    return OR_OK;
}

/* EriSchwarzBounds (eri.cpp:587-620): Q[a][b] = Q[b][a] = the next double
 * above sqrt(max |(a_i b_j | b_j a_i)|) * (1 + 64 eps), shell pairs a <= b. */
int or_schwarz_bounds(const or_basis* b, const double field[6], double* out) {
    const int ns = b->nshell;
    shell_view* sv;
    int st = make_all_shells(b, &sv);
    if (st) return st;
    workspace w;
    memset(&w, 0, sizeof w);
    cplx* block = NULL;
    size_t cap = 0;
    const double safety = 1.0 + 64.0 * 2.220446049250313e-16;
    for (int a = 0; a < ns && !st; ++a) {
        for (int c = a; c < ns; ++c) {
            const uint32_t q[4] = {(uint32_t)a, (uint32_t)c, (uint32_t)c, (uint32_t)a};
            const size_t size = or_quartet_size(b, q);
            block = (cplx*)grow(block, &cap, size, sizeof(cplx));
            const shell_view* sh[4] = {&sv[a], &sv[c], &sv[c], &sv[a]};
            st = compute_eri_shell(sh, field, block, &w);
            if (st) break;
            const size_t na = shell_ao_count(b, a), nb = shell_ao_count(b, c);
            double m = 0.0;
            for (size_t i = 0; i < na; ++i)
                for (size_t j = 0; j < nb; ++j) {
                    const cplx v = block[((i * nb + j) * nb + j) * na + i];
                    const double r = cabs(v);
                    if (r > m) m = r;
                }
            const double bound = nextafter(sqrt(m) * safety, INFINITY);
            out[(size_t)a * ns + c] = bound;
            out[(size_t)c * ns + a] = bound;
        }
    }
    free(block);
    ws_free(&w);
    free_all_shells(b, sv);
    return st;
}

/* ------------------------------------------------------------------------ */
/* Centre derivatives (derivatives.cpp)                                     */
/* ------------------------------------------------------------------------ */

/* derivatives.cpp:43-51: (B x e_axis) */
static void field_cross_axis(const double field[6], int axis, double out[3]) {
    const double e[3] = {axis == 0, axis == 1, axis == 2};
    out[0] = field[1] * e[2] - field[2] * e[1];
    out[1] = field[2] * e[0] - field[0] * e[2];
    out[2] = field[0] * e[1] - field[1] * e[0];
}

/* derivatives.cpp:54-84 center_response: d/dA_axis of the raw primitive
 * quartet P with respect to the centre of slot `slot` (slots 0 and 2 are the
 * conjugated functions, London phase sign +1/2; slots 1 and 3 take -1/2). */
static int center_response(const prim_t P[4], int slot, int axis, const double field[6], const pair_data* ab,
                           const pair_data* cd, int real_zero, workspace* w, cplx* out) {
    prim_t Q[4];
    memcpy(Q, P, sizeof Q);
    cplx result = 0.0, v;
    int st;
    const int a = P[slot].ang[axis];
    if (a != 0) {
        Q[slot].ang[axis] = a - 1;
        if ((st = eri_kernel(&Q[0], &Q[1], &Q[2], &Q[3], ab, cd, real_zero, w, &v))) return st;
        result -= (double)a * v;
    }
    Q[slot].ang[axis] = a + 1;
    if ((st = eri_kernel(&Q[0], &Q[1], &Q[2], &Q[3], ab, cd, real_zero, w, &v))) return st;
    result += 2.0 * P[slot].exponent * v;
    Q[slot].ang[axis] = a;
    double phase[3];
    field_cross_axis(field, axis, phase);
    const double sign = slot % 2 == 0 ? 0.5 : -0.5;
    cplx base;
    if ((st = eri_kernel(&P[0], &P[1], &P[2], &P[3], ab, cd, real_zero, w, &base))) return st;
    for (int k = 0; k < 3; ++k) {
        const double coef = sign * phase[k];
        if (coef == 0.0) continue;
        Q[slot].ang[k] += 1;
        st = eri_kernel(&Q[0], &Q[1], &Q[2], &Q[3], ab, cd, real_zero, w, &v);
        Q[slot].ang[k] -= 1;
        if (st) return st;
        const cplx cv = v + P[slot].center[k] * base;
        result += CMPLX(-coef * cimag(cv), coef * creal(cv));
    }
    *out = result;
    return OR_OK;
}

/* derivatives.cpp:556-660 compute_quartet_block with the evaluator of
 * derivatives.cpp:232-241,836-842: block (4, 3, na, nb, nc, nd). Each raw
 * primitive carries its normalization in the coefficient (raw_primitive,
 * derivatives.cpp:30-37); contraction weights multiply the 12 responses. */
static int compute_derivative_shell(const shell_view* sh[4], const double field[6], cplx* out, workspace* w) {
    const shell_view *a = sh[0], *b = sh[1], *c = sh[2], *d = sh[3];
    const size_t na = (size_t)a->ncontr * a->ncart, nb = (size_t)b->ncontr * b->ncart;
    const size_t nc = (size_t)c->ncontr * c->ncart, nd = (size_t)d->ncontr * d->ncart;
    const size_t size = na * nb * nc * nd;
    for (size_t i = 0; i < 12 * size; ++i) out[i] = 0.0;
    const int real_zero = use_real_path(field);
    for (int pa = 0; pa < a->nprim; ++pa)
        for (int pb = 0; pb < b->nprim; ++pb) {
            pair_data ab;
            gaussian_product(a->exps[pa], a->center, b->exps[pb], b->center, field, &ab);
            for (int pc = 0; pc < c->nprim; ++pc)
                for (int pd = 0; pd < d->nprim; ++pd) {
                    pair_data cdp;
                    gaussian_product(c->exps[pc], c->center, d->exps[pd], d->center, field, &cdp);
                    for (int ca = 0; ca < a->ncart; ++ca)
                        for (int cb = 0; cb < b->ncart; ++cb)
                            for (int cc = 0; cc < c->ncart; ++cc)
                                for (int cq = 0; cq < d->ncart; ++cq) {
                                    prim_t P[4];
                                    const shell_view* s4[4] = {a, b, c, d};
                                    const int pp[4] = {pa, pb, pc, pd}, cp[4] = {ca, cb, cc, cq};
                                    for (int k = 0; k < 4; ++k) {
                                        make_prim(s4[k], pp[k], cp[k], &P[k]);
                                        P[k].coefficient = s4[k]->norms[(size_t)cp[k] * s4[k]->nprim + pp[k]];
                                    }
                                    cplx values[12];
                                    for (int slot = 0; slot < 4; ++slot)
                                        for (int axis = 0; axis < 3; ++axis) {
                                            const int st = center_response(P, slot, axis, field, &ab, &cdp, real_zero, w,
                                                                           &values[slot * 3 + axis]);
                                            if (st) return st;
                                        }
                                    for (int xa = 0; xa < a->ncontr; ++xa) {
                                        const size_t ia = (size_t)xa * a->ncart + ca;
                                        for (int xb = 0; xb < b->ncontr; ++xb) {
                                            const size_t ib = (size_t)xb * b->ncart + cb;
                                            for (int xc = 0; xc < c->ncontr; ++xc) {
                                                const size_t ic = (size_t)xc * c->ncart + cc;
                                                for (int xd = 0; xd < d->ncontr; ++xd) {
                                                    const size_t id = (size_t)xd * d->ncart + cq;
                                                    const double weight = a->coefs[(size_t)xa * a->nprim + pa] *
                                                                          b->coefs[(size_t)xb * b->nprim + pb] *
                                                                          c->coefs[(size_t)xc * c->nprim + pc] *
                                                                          d->coefs[(size_t)xd * d->nprim + pd];
                                                    const size_t e = ((ia * nb + ib) * nc + ic) * nd + id;
                                                    for (int k = 0; k < 12; ++k) out[k * size + e] += weight * values[k];
                                                }
                                            }
                                        }
                                    }
                                }
                }
        }
    return OR_OK;
}

/* compute_eri_derivative_blocks (derivatives.cpp:1389-1415): offsets holds
 * n + 1 entries starting at 0, block i spans offsets[i]..offsets[i+1] and
 * has 12 x the quartet size; offsets[n] must equal out_len. */
int or_eri_derivative_blocks(const or_basis* b, const uint32_t* quartets, size_t n, const double field[6],
                             const uint64_t* offsets, double* out, size_t out_len, int threads) {
    if (threads < 1) return fail(OR_INVALID_ARGUMENT, "ERI thread count must be positive");
    if (offsets[0] != 0) return fail(OR_LENGTH_ERROR, "derivative batch offsets have an incorrect size or start");
    for (size_t i = 0; i < n; ++i) {
        for (int k = 0; k < 4; ++k)
            if (quartets[4 * i + k] >= (uint32_t)b->nshell)
                return fail(OR_OUT_OF_RANGE, "derivative shell quartet index is out of range");
        const size_t size = or_quartet_size(b, quartets + 4 * i);
        if (offsets[i] > out_len || offsets[i + 1] < offsets[i] || offsets[i + 1] - offsets[i] != 12 * size)
            return fail(OR_LENGTH_ERROR, "derivative batch offsets do not match block sizes");
    }
    if (offsets[n] != out_len) return fail(OR_LENGTH_ERROR, "derivative batch output has an incorrect size");
    shell_view* sv;
    int st = make_all_shells(b, &sv);
    if (st) return st;
    int first_error = OR_OK;
    char first_msg[512] = {0};
#pragma omp parallel num_threads(threads)
    {
        workspace w;
        memset(&w, 0, sizeof w);
#pragma omp for schedule(dynamic, 1)
        for (long long i = 0; i < (long long)n; ++i) {
            int failed;
#pragma omp atomic read
            failed = first_error;
            if (failed) continue;
            const uint32_t* q = quartets + 4 * i;
            const shell_view* sh[4] = {&sv[q[0]], &sv[q[1]], &sv[q[2]], &sv[q[3]]};
            const int s = compute_derivative_shell(sh, field, (cplx*)(out + 2 * offsets[i]), &w);
            if (s) {
#pragma omp critical
                {
                    if (!first_error) {
                        first_error = s;
                        memcpy(first_msg, g_msg, sizeof first_msg);
                    }
                }
            }
        }
        ws_free(&w);
    }
    free_all_shells(b, sv);
    if (first_error) {
        memcpy(g_msg, first_msg, sizeof g_msg);
        return first_error;
    }
    return OR_OK;
}

/* EriDerivativeSchwarzBounds (derivatives.cpp:1307-1372): out[a*ns+b] bounds
 * every centre derivative of the pair density (a b| over slots 0 and 1 by a
 * triangle/Schwarz sum of sqrt|(right left|left right)| of the shifted raw
 * primitives, contracted with |coefficients|, padded by 128 eps and one ulp. */
int or_eri_derivative_pair_bounds(const or_basis* bs, const double field[6], double* out) {
    const int ns = bs->nshell;
    shell_view* sv;
    int st = make_all_shells(bs, &sv);
    if (st) return st;
    workspace w;
    memset(&w, 0, sizeof w);
    const int real_zero = use_real_path(field);
    /* norm(left, right) of derivatives.cpp:1325-1333 */
#define DERIV_NORM(L, R, OUT)                                                               \
    do {                                                                                    \
        pair_data p1, p2;                                                                   \
        gaussian_product((R).exponent, (R).center, (L).exponent, (L).center, field, &p1);   \
        gaussian_product((L).exponent, (L).center, (R).exponent, (R).center, field, &p2);   \
        cplx v_;                                                                            \
        st = eri_kernel(&(R), &(L), &(L), &(R), &p1, &p2, real_zero, &w, &v_);              \
        (OUT) = sqrt(cabs(v_));                                                             \
    } while (0)
    for (int ia = 0; ia < ns && !st; ++ia)
        for (int ib = 0; ib < ns && !st; ++ib) {
            const shell_view *a = &sv[ia], *b = &sv[ib];
            double maximum = 0.0;
            double* contracted = calloc((size_t)6 * a->ncontr * b->ncontr, sizeof(double));
            for (int ca = 0; ca < a->ncart && !st; ++ca)
                for (int cb = 0; cb < b->ncart && !st; ++cb) {
                    memset(contracted, 0, sizeof(double) * 6 * (size_t)a->ncontr * b->ncontr);
                    for (int pa = 0; pa < a->nprim && !st; ++pa)
                        for (int pb = 0; pb < b->nprim && !st; ++pb) {
                            prim_t ga, gb;
                            make_prim(a, pa, ca, &ga);
                            make_prim(b, pb, cb, &gb);
                            ga.coefficient = a->norms[(size_t)ca * a->nprim + pa];
                            gb.coefficient = b->norms[(size_t)cb * b->nprim + pb];
                            double pbnd[6];
                            for (int slot = 0; slot < 2 && !st; ++slot) {
                                const prim_t raw = slot == 0 ? ga : gb;
                                for (int axis = 0; axis < 3 && !st; ++axis) {
                                    prim_t ch = raw;
                                    double nv;
                                    ch.ang[axis] += 1;
                                    if (slot == 0) DERIV_NORM(ch, gb, nv); else DERIV_NORM(ga, ch, nv);
                                    double bound = 2.0 * raw.exponent * nv;
                                    if (raw.ang[axis] != 0) {
                                        ch = raw;
                                        ch.ang[axis] -= 1;
                                        if (slot == 0) DERIV_NORM(ch, gb, nv); else DERIV_NORM(ga, ch, nv);
                                        bound += (double)raw.ang[axis] * nv;
                                    }
                                    double phase[3];
                                    field_cross_axis(field, axis, phase);
                                    for (int k = 0; k < 3 && !st; ++k) {
                                        const double weight = 0.5 * fabs(phase[k]);
                                        if (weight == 0.0) continue;
                                        double up, same;
                                        ch = raw;
                                        ch.ang[k] += 1;
                                        if (slot == 0) DERIV_NORM(ch, gb, up); else DERIV_NORM(ga, ch, up);
                                        DERIV_NORM(ga, gb, same);
                                        bound += weight * (up + fabs(raw.center[k]) * same);
                                    }
                                    pbnd[slot * 3 + axis] = bound;
                                }
                            }
                            for (int xa = 0; xa < a->ncontr; ++xa)
                                for (int xb = 0; xb < b->ncontr; ++xb) {
                                    const double weight = fabs(a->coefs[(size_t)xa * a->nprim + pa] *
                                                               b->coefs[(size_t)xb * b->nprim + pb]);
                                    for (int k = 0; k < 6; ++k)
                                        contracted[((size_t)xa * b->ncontr + xb) * 6 + k] += weight * pbnd[k];
                                }
                        }
                    for (size_t k = 0; k < (size_t)6 * a->ncontr * b->ncontr; ++k)
                        if (contracted[k] > maximum) maximum = contracted[k];
                }
            free(contracted);
            out[(size_t)ia * ns + ib] =
                nextafter(maximum * (1.0 + 128.0 * 2.220446049250313e-16), INFINITY);
        }
#undef DERIV_NORM
    ws_free(&w);
    free_all_shells(bs, sv);
    return st;
}

/* eri.cpp:414-429 */
int or_canonicalize(const uint32_t in[4], uint32_t out[4]) {
    const uint32_t cand[4][4] = {{in[0], in[1], in[2], in[3]},
                                 {in[2], in[3], in[0], in[1]},
                                 {in[1], in[0], in[3], in[2]},
                                 {in[3], in[2], in[1], in[0]}};
    int sel = 0;
    for (int i = 1; i < 4; ++i) {
        int less = 0;
        for (int k = 0; k < 4; ++k) {
            if (cand[i][k] != cand[sel][k]) {
                less = cand[i][k] < cand[sel][k];
                break;
            }
        }
        if (less) sel = i;
    }
    for (int k = 0; k < 4; ++k) out[k] = cand[sel][k];
    return sel >= 2;
}

/* eri.cpp:777-847 */
int or_eri_tensor(const or_basis* b, const double field[6], double* out, int threads) {
    const int ns = b->nshell;
    size_t nao = 0;
    size_t* aoff = malloc(sizeof(size_t) * (size_t)(ns + 1));
    for (int s = 0; s < ns; ++s) {
        aoff[s] = nao;
        nao += shell_ao_count(b, s);
    }
    size_t nq = 0, cap = 64;
    uint32_t* qs = malloc(sizeof(uint32_t) * 4 * cap);
    for (uint32_t a = 0; a < (uint32_t)ns; ++a)
        for (uint32_t bb = 0; bb < (uint32_t)ns; ++bb)
            for (uint32_t c = 0; c < (uint32_t)ns; ++c)
                for (uint32_t d = 0; d < (uint32_t)ns; ++d) {
                    const uint32_t q[4] = {a, bb, c, d};
                    uint32_t cq[4];
                    or_canonicalize(q, cq);
                    if (memcmp(q, cq, sizeof q) != 0) continue;
                    if (nq == cap) {
                        cap *= 2;
                        qs = realloc(qs, sizeof(uint32_t) * 4 * cap);
                    }
                    memcpy(qs + 4 * nq, q, sizeof q);
                    ++nq;
                }
    uint64_t* offs = malloc(sizeof(uint64_t) * (nq + 1));
    uint64_t total = 0;
    for (size_t i = 0; i < nq; ++i) {
        offs[i] = total;
        total += or_quartet_size(b, qs + 4 * i);
    }
    double* blocks = malloc(sizeof(double) * 2 * (total ? total : 1));
    int st = or_eri_batch(b, qs, nq, field, offs, blocks, threads);
    if (!st) {
        const size_t n1 = nao, n2 = nao * nao, n3 = n2 * nao;
        memset(out, 0, sizeof(double) * 2 * n3 * nao);
        for (size_t i = 0; i < nq; ++i) {
            const uint32_t* q = qs + 4 * i;
            const size_t sa = shell_ao_count(b, (int)q[0]), sb = shell_ao_count(b, (int)q[1]);
            const size_t sc = shell_ao_count(b, (int)q[2]), sd = shell_ao_count(b, (int)q[3]);
            const double* v = blocks + 2 * offs[i];
            for (size_t ia = 0; ia < sa; ++ia)
                for (size_t ib = 0; ib < sb; ++ib)
                    for (size_t ic = 0; ic < sc; ++ic)
                        for (size_t id = 0; id < sd; ++id) {
                            const size_t k = ((ia * sb + ib) * sc + ic) * sd + id;
                            const double re = v[2 * k], im = v[2 * k + 1];
                            const size_t ga = aoff[q[0]] + ia, gb = aoff[q[1]] + ib;
                            const size_t gc = aoff[q[2]] + ic, gd = aoff[q[3]] + id;
                            size_t t;
                            t = ga * n3 + gb * n2 + gc * n1 + gd;
                            out[2 * t] = re, out[2 * t + 1] = im;
                            t = gc * n3 + gd * n2 + ga * n1 + gb;
                            out[2 * t] = re, out[2 * t + 1] = im;
                            t = gb * n3 + ga * n2 + gd * n1 + gc;
                            out[2 * t] = re, out[2 * t + 1] = -im;
                            t = gd * n3 + gc * n2 + gb * n1 + ga;
                            out[2 * t] = re, out[2 * t + 1] = -im;
                        }
        }
    }
    free(blocks);
    free(offs);
    free(qs);
    free(aoff);
    return st;
}

/* ------------------------------------------------------------------------ */
/* cholesky.cpp                                                             */
/* ------------------------------------------------------------------------ */

/* cholesky.cpp:33-70 SelectedQuartet::evaluate for one AO element. */
static int selected_eval(const shell_view* sh[4], const size_t aos[4], const double field[6], workspace* w,
                         cplx* result) {
    int comp[4], contr[4];
    for (int s = 0; s < 4; ++s) {
        comp[s] = (int)(aos[s] % (size_t)sh[s]->ncart);
        contr[s] = (int)(aos[s] / (size_t)sh[s]->ncart);
    }
    const shell_view *a = sh[0], *b = sh[1], *c = sh[2], *d = sh[3];
    const int real_zero = use_real_path(field);
    cplx r = 0.0;
    for (int pa = 0; pa < a->nprim; ++pa)
        for (int pb = 0; pb < b->nprim; ++pb) {
            pair_data ab;
            gaussian_product(a->exps[pa], a->center, b->exps[pb], b->center, field, &ab);
            for (int pc = 0; pc < c->nprim; ++pc)
                for (int pd = 0; pd < d->nprim; ++pd) {
                    pair_data cd;
                    gaussian_product(c->exps[pc], c->center, d->exps[pd], d->center, field, &cd);
                    prim_t P[4];
                    make_prim(a, pa, comp[0], &P[0]);
                    make_prim(b, pb, comp[1], &P[1]);
                    make_prim(c, pc, comp[2], &P[2]);
                    make_prim(d, pd, comp[3], &P[3]);
                    cplx pv;
                    const int st = eri_kernel(&P[0], &P[1], &P[2], &P[3], &ab, &cd, real_zero, w, &pv);
                    if (st) return st;
                    pv *= a->norms[(size_t)comp[0] * a->nprim + pa];
                    pv *= b->norms[(size_t)comp[1] * b->nprim + pb];
                    pv *= c->norms[(size_t)comp[2] * c->nprim + pc];
                    pv *= d->norms[(size_t)comp[3] * d->nprim + pd];
                    r += a->coefs[(size_t)contr[0] * a->nprim + pa] * b->coefs[(size_t)contr[1] * b->nprim + pb] *
                         c->coefs[(size_t)contr[2] * c->nprim + pc] * d->coefs[(size_t)contr[3] * d->nprim + pd] * pv;
                }
        }
    *result = r;
    return OR_OK;
}

/* cholesky.cpp:285-314 */
int or_cholesky_diagonal(const or_basis* b, const uint64_t* pairs, size_t npairs, const double field[6],
                         double* out, int threads) {
    if (threads < 1) return fail(OR_INVALID_ARGUMENT, "threads must be a positive integer <= 2147483647");
    const int ns = b->nshell;
    size_t* aoff = malloc(sizeof(size_t) * (size_t)(ns + 1));
    size_t n = 0;
    for (int s = 0; s < ns; ++s) {
        aoff[s] = n;
        n += shell_ao_count(b, s);
    }
    memset(out, 0, sizeof(double) * n * n);
    char* seen = calloc((size_t)ns * (size_t)ns + 1, 1);
    for (size_t i = 0; i < npairs; ++i) {
        const uint64_t a = pairs[2 * i], bb = pairs[2 * i + 1];
        if (a >= (uint64_t)ns || bb >= (uint64_t)ns || a > bb) {
            free(seen);
            free(aoff);
            return fail(OR_OUT_OF_RANGE, "Cholesky diagonal shell pair must be valid and ordered");
        }
        if (seen[a * (uint64_t)ns + bb]) {
            free(seen);
            free(aoff);
            return fail(OR_INVALID_ARGUMENT, "duplicate Cholesky diagonal shell pair");
        }
        seen[a * (uint64_t)ns + bb] = 1;
    }
    free(seen);
    shell_view* sv;
    int st = make_all_shells(b, &sv);
    if (st) {
        free(aoff);
        return st;
    }
    int first_error = OR_OK;
    char first_msg[512] = {0};
#pragma omp parallel num_threads(threads)
    {
        workspace w;
        memset(&w, 0, sizeof w);
#pragma omp for schedule(dynamic, 1)
        for (long long i = 0; i < (long long)npairs; ++i) {
            const int a = (int)pairs[2 * i], bb = (int)pairs[2 * i + 1];
            const shell_view* sh[4] = {&sv[a], &sv[bb], &sv[bb], &sv[a]};
            const size_t na = shell_ao_count(b, a), nb = shell_ao_count(b, bb);
            for (size_t ii = 0; ii < na; ++ii)
                for (size_t jj = 0; jj < nb; ++jj) {
                    const size_t aos[4] = {ii, jj, jj, ii};
                    cplx v;
                    const int s = selected_eval(sh, aos, field, &w, &v);
                    if (s) {
#pragma omp critical
                        {
                            if (!first_error) {
                                first_error = s;
                                memcpy(first_msg, g_msg, sizeof first_msg);
                            }
                        }
                        continue;
                    }
                    const size_t mu = aoff[a] + ii, nu = aoff[bb] + jj;
                    out[nu * n + mu] = creal(v);
                    if (a != bb) out[mu * n + nu] = creal(v);
                }
        }
        ws_free(&w);
    }
    free_all_shells(b, sv);
    free(aoff);
    if (first_error) {
        memcpy(g_msg, first_msg, sizeof g_msg);
        return first_error;
    }
    return OR_OK;
}

/* cholesky.cpp:203-282 (per-element evaluation; values equal contracted_columns). */
int or_cholesky_columns(const or_basis* b, const uint64_t* pivots, size_t npiv, const double field[6],
                        double* out, int real_output, int threads) {
    if (threads < 1) return fail(OR_INVALID_ARGUMENT, "threads must be a positive integer <= 2147483647");
    const int ns = b->nshell;
    const int real = use_real_path(field);
    if (real_output && !real) return fail(OR_INVALID_ARGUMENT, "real Cholesky columns require zero phase field");
    if (npiv == 0) return OR_OK;
    size_t* aoff = malloc(sizeof(size_t) * (size_t)(ns + 1));
    size_t n = 0;
    for (int s = 0; s < ns; ++s) {
        aoff[s] = n;
        n += shell_ao_count(b, s);
    }
    int* owner = malloc(sizeof(int) * (n ? n : 1));
    for (int s = 0; s < ns; ++s)
        for (size_t k = 0; k < shell_ao_count(b, s); ++k) owner[aoff[s] + k] = s;
    for (size_t i = 0; i < npiv; ++i)
        if (pivots[2 * i] >= n || pivots[2 * i + 1] >= n) {
            free(owner);
            free(aoff);
            return fail(OR_OUT_OF_RANGE, "Cholesky pivot AO index is out of range");
        }
    const int sa = owner[pivots[0]], sb = owner[pivots[1]];
    for (size_t i = 0; i < npiv; ++i)
        if (owner[pivots[2 * i]] != sa || owner[pivots[2 * i + 1]] != sb) {
            free(owner);
            free(aoff);
            return fail(OR_INVALID_ARGUMENT, "Cholesky candidate pivots must share a shell pair");
        }
    shell_view* sv;
    int st = make_all_shells(b, &sv);
    if (st) {
        free(owner);
        free(aoff);
        return st;
    }
    /* At exact zero field only b >= a is evaluated and mirrored (cholesky.cpp:224,258). */
    const long long npairs = (long long)ns * ns;
    int first_error = OR_OK;
    char first_msg[512] = {0};
#pragma omp parallel num_threads(threads)
    {
        workspace w;
        memset(&w, 0, sizeof w);
#pragma omp for schedule(dynamic, 1)
        for (long long g = 0; g < npairs; ++g) {
            const int a = (int)(g / ns), bb = (int)(g % ns);
            if (real && bb < a) continue;
            const shell_view* sh[4] = {&sv[a], &sv[bb], &sv[sb], &sv[sa]};
            const size_t na = shell_ao_count(b, a), nb = shell_ao_count(b, bb);
            for (size_t cand = 0; cand < npiv; ++cand) {
                const size_t cl = pivots[2 * cand + 1] - aoff[sb], dl = pivots[2 * cand] - aoff[sa];
                for (size_t i = 0; i < na; ++i)
                    for (size_t j = 0; j < nb; ++j) {
                        const size_t aos[4] = {i, j, cl, dl};
                        cplx v;
                        int s = selected_eval(sh, aos, field, &w, &v);
                        if (!s && real_output && a == bb) {
                            const size_t aos2[4] = {j, i, cl, dl};
                            cplx v2;
                            s = selected_eval(sh, aos2, field, &w, &v2);
                            if (!s) v = (creal(v) + creal(v2)) * 0.5;
                        }
                        if (s) {
#pragma omp critical
                            {
                                if (!first_error) {
                                    first_error = s;
                                    memcpy(first_msg, g_msg, sizeof first_msg);
                                }
                            }
                            continue;
                        }
                        const size_t mu = aoff[a] + i, nu = aoff[bb] + j;
                        if (real_output) {
                            out[(cand * n + mu) * n + nu] = creal(v);
                            if (a != bb) out[(cand * n + nu) * n + mu] = creal(v);
                        } else {
                            size_t k = (cand * n + mu) * n + nu;
                            out[2 * k] = creal(v);
                            out[2 * k + 1] = cimag(v);
                            if (real && a != bb) {
                                k = (cand * n + nu) * n + mu;
                                out[2 * k] = creal(v);
                                out[2 * k + 1] = cimag(v);
                            }
                        }
                    }
            }
        }
        ws_free(&w);
    }
    free_all_shells(b, sv);
    free(owner);
    free(aoff);
    if (first_error) {
        memcpy(g_msg, first_msg, sizeof g_msg);
        return first_error;
    }
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* nuclear.cpp: nuclear attraction with London orbitals                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    int side;
    double wave[3];
    cplx disp[3];
    cplx* table;          /* [n][t][u][v] */
    unsigned char* ready;
} nuc_aux;

/* nuclear.cpp:106-141: the four-index auxiliary with the pair's wave term,
 * memoized, in the reference's recursion order. */
static cplx nuc_auxiliary(nuc_aux* a, int t, int u, int v, int n) {
    const size_t side = (size_t)a->side;
    const size_t idx = (((size_t)n * side + t) * side + u) * side + v;
    if (a->ready[idx]) return a->table[idx];
    cplx value;
    if (t > 0) {
        value = CMPLX(0.0, -a->wave[0]) * nuc_auxiliary(a, t - 1, u, v, n) +
                a->disp[0] * nuc_auxiliary(a, t - 1, u, v, n + 1);
        if (t > 1) value += (double)(t - 1) * nuc_auxiliary(a, t - 2, u, v, n + 1);
    } else if (u > 0) {
        value = CMPLX(0.0, -a->wave[1]) * nuc_auxiliary(a, t, u - 1, v, n) +
                a->disp[1] * nuc_auxiliary(a, t, u - 1, v, n + 1);
        if (u > 1) value += (double)(u - 1) * nuc_auxiliary(a, t, u - 2, v, n + 1);
    } else {
        value = CMPLX(0.0, -a->wave[2]) * nuc_auxiliary(a, t, u, v - 1, n) +
                a->disp[2] * nuc_auxiliary(a, t, u, v - 1, n + 1);
        if (v > 1) value += (double)(v - 1) * nuc_auxiliary(a, t, u, v - 2, n + 1);
    }
    a->table[idx] = value;
    a->ready[idx] = 1;
    return value;
}

/* nuclear.cpp:24-153 NuclearKernel::compute for one nucleus, unnormalized
 * primitives with unit coefficients (as compute_nuclear_attraction calls it). */
static int nuclear_kernel(const prim_t* bra, const prim_t* ket, double charge, const double C[3],
                          const double field[6], cplx* value) {
    pair_data pair;
    gaussian_product(bra->exponent, bra->center, ket->exponent, ket->center, field, &pair);
    const int L = bra->ang[0] + bra->ang[1] + bra->ang[2] + ket->ang[0] + ket->ang[1] + ket->ang[2];
    if (L > 32) return fail(OR_INVALID_ARGUMENT, "nuclear attraction supports combined angular order through 32");
    double E[3][2 * BOYS_MAX_ORDER + 4];
    for (int axis = 0; axis < 3; ++axis)
        hermite_coefficients(bra->ang[axis], ket->ang[axis], bra->center[axis], ket->center[axis], pair.P[axis],
                             pair.exponent, E[axis]);
    nuc_aux a;
    a.side = L + 1;
    for (int i = 0; i < 3; ++i) {
        a.wave[i] = pair.q[i];
        a.disp[i] = pair.Pc[i] - C[i];
    }
    const cplx arg = pair.exponent * (a.disp[0] * a.disp[0] + a.disp[1] * a.disp[1] + a.disp[2] * a.disp[2]);
    const int scaled = creal(arg) < 0.0;
    cplx F[BOYS_MAX_ORDER + 1];
    int region;
    double err;
    size_t segs;
    int st = compute_boys(arg, L, scaled, F, &region, &err, &segs);
    if (st) return st;
    cplx common = pair.london;
    if (scaled) {
        const double d[3] = {pair.P[0] - C[0], pair.P[1] - C[1], pair.P[2] - C[2]};
        common = cexp(CMPLX(-pair.exponent * dot3(d, d), -dot3(pair.q, C)));
    }
    const size_t side = (size_t)a.side, cells = side * side * side * side;
    a.table = malloc(sizeof(cplx) * cells);
    a.ready = calloc(cells, 1);
    if (!a.table || !a.ready) {
        free(a.table);
        free(a.ready);
        return fail(OR_RUNTIME_ERROR, "out of memory");
    }
    double radial = 1.0;
    for (int n = 0; n <= L; ++n) {
        const size_t idx = (size_t)n * side * side * side;
        a.table[idx] = common * radial * F[n];
        a.ready[idx] = 1;
        radial *= -2.0 * pair.exponent;
    }
    cplx contraction = 0.0;
    const int mx = bra->ang[0] + ket->ang[0], my = bra->ang[1] + ket->ang[1], mz = bra->ang[2] + ket->ang[2];
    for (int t = 0; t <= mx; ++t)
        for (int u = 0; u <= my; ++u)
            for (int v = 0; v <= mz; ++v)
                contraction += E[0][t] * E[1][u] * E[2][v] * nuc_auxiliary(&a, t, u, v, 0);
    free(a.table);
    free(a.ready);
    *value = -charge * (2.0 * PI / pair.exponent) * pair.ordinary * contraction;
    return OR_OK;
}

/* nuclear.cpp:209-280 compute_nuclear_attraction and 282-312
 * compute_nuclear_attraction_matrix: every ordered shell pair's block. */
int or_nuclear_matrix(const or_basis* b, int nnuc, const double* charges, const double* centers,
                      const double field[6], double* out) {
    shell_view* sh = NULL;
    int st = make_all_shells(b, &sh);
    if (st) return st;
    size_t nao = 0;
    size_t* off = malloc(sizeof(size_t) * (size_t)(b->nshell + 1));
    if (!off) {
        free_all_shells(b, sh);
        return fail(OR_RUNTIME_ERROR, "out of memory");
    }
    for (int s = 0; s < b->nshell; ++s) {
        off[s] = nao;
        nao += shell_ao_count(b, s);
    }
    cplx* V = (cplx*)out;
    for (size_t i = 0; i < nao * nao; ++i) V[i] = 0.0;
    for (int sa = 0; sa < b->nshell && !st; ++sa) {
        for (int sb = 0; sb < b->nshell && !st; ++sb) {
            const shell_view *A = &sh[sa], *B = &sh[sb];
            for (int pa = 0; pa < A->nprim && !st; ++pa)
                for (int pb = 0; pb < B->nprim && !st; ++pb)
                    for (int ca = 0; ca < A->ncart && !st; ++ca)
                        for (int cb = 0; cb < B->ncart && !st; ++cb) {
                            prim_t bra, ket;
                            make_prim(A, pa, ca, &bra);
                            make_prim(B, pb, cb, &ket);
                            cplx value = 0.0;
                            for (int c = 0; c < nnuc && !st; ++c) {
                                cplx v;
                                st = nuclear_kernel(&bra, &ket, charges[c], centers + 3 * c, field, &v);
                                if (!st) value += v;
                            }
                            value *= A->norms[(size_t)ca * A->nprim + pa] * B->norms[(size_t)cb * B->nprim + pb];
                            for (int xa = 0; xa < A->ncontr; ++xa)
                                for (int xb = 0; xb < B->ncontr; ++xb) {
                                    const size_t mu = off[sa] + (size_t)xa * A->ncart + ca;
                                    const size_t nu = off[sb] + (size_t)xb * B->ncart + cb;
                                    V[mu * nao + nu] += A->coefs[(size_t)xa * A->nprim + pa] *
                                                        B->coefs[(size_t)xb * B->nprim + pb] * value;
                                }
                        }
        }
    }
    free(off);
    free_all_shells(b, sh);
    return st;
}
