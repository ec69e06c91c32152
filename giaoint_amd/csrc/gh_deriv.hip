// ERI centre derivatives: combination of shifted-momentum integrals.
//
// GIAOINT differentiates each raw primitive quartet (derivatives.cpp:54-84,
// 232-241): for the centre of slot k and axis x,
//   d/dA_x = -a_x (a - 1_x| + 2 alpha (a + 1_x|
//            + i s_k sum_y (B x e_x)_y [ (a + 1_y| + A_y (a| ],
// with s_k = +1/2 for the conjugated slots 0 and 2 and -1/2 for slots 1 and
// 3, the primitive normalization of the unshifted component kept fixed, and
// the contraction weights applied after the combination
// (derivatives.cpp:556-660). The weights 2 alpha and the radial
// normalization ratio depend on the primitive only, so the host folds them
// into the contraction coefficients of shifted shells (l + 1 with rows
// c sqrt(alpha) then c / (2 sqrt(alpha)), l - 1 with c 2 sqrt(alpha)); the
// ordinary ERI path evaluates those contracted quartets, and this kernel
// forms the twelve derivative blocks from them, replacing the angular part
// of the shifted component's normalization by the unshifted one's:
// N(c) / N(c + 1_x) = sqrt(2 a_x + 1), N(c) / N(c - 1_x) = 1 / sqrt(2 a_x - 1).
#include "gh_internal.h"

namespace gh {

namespace {

__device__ inline int ncart_of(int l) { return (l + 1) * (l + 2) / 2; }

// Position of (ax, ay, az) in the reference's Cartesian order (x-major,
// then y, descending).
__device__ inline int comp_index(int ax, int ay, int az) {
    const int i = ay + az;  // = l - ax
    return i * (i + 1) / 2 + az;
}

__device__ inline void comp_powers(int l, int c, int p[3]) {
    int ax = l, first = 0;
    while (c >= first + (l - ax + 1)) {
        first += l - ax + 1;
        --ax;
    }
    const int az = c - first;
    p[0] = ax;
    p[1] = l - ax - az;
    p[2] = az;
}

__global__ void __launch_bounds__(256) deriv_assemble_kernel(const DerivItem* __restrict__ items,
                                                             const int* __restrict__ tile_item, const zc* __restrict__ t,
                                                             zc* __restrict__ out, double bx, double by, double bz,
                                                             int phase) {
    const int item = tile_item[blockIdx.x];
    const DerivItem& it = items[item];
    const long long e = (long long)(blockIdx.x - it.first_tile) * 256 + threadIdx.x;
    if (e >= it.size) return;
    int ao[4];
    long long r = e;
    for (int k = 3; k >= 0; --k) {
        ao[k] = (int)(r % it.n[k]);
        r /= it.n[k];
    }
    const zc base = t[it.base + e];
    const double B[3] = {bx, by, bz};
    zc* dst = out + it.out;
    for (int k = 0; k < 4; ++k) {
        const int l = it.l[k], nc = ncart_of(l), row = ao[k] / nc;
        int pw[3];
        comp_powers(l, ao[k] - row * nc, pw);
        // flat index in a block whose slot k has `nk` AOs, with AO `aok` there
        auto at = [&](int nk, int aok) {
            long long idx = 0;
            for (int j = 0; j < 4; ++j) idx = idx * (j == k ? nk : it.n[j]) + (j == k ? aok : ao[j]);
            return idx;
        };
        const int ncp = ncart_of(l + 1), ncm = l > 0 ? ncart_of(l - 1) : 0;
        const int np = it.prow[k] * ncp, nm = it.krow[k] * ncm;
        const double sign = (k % 2 == 0) ? 0.5 : -0.5;
        for (int x = 0; x < 3; ++x) {
            int q[3] = {pw[0], pw[1], pw[2]};
            double re = 0.0, im = 0.0;
            const int a = pw[x];
            if (a > 0) {
                q[x] = a - 1;
                const zc v = t[it.minus[k] + at(nm, row * ncm + comp_index(q[0], q[1], q[2]))];
                const double w = (double)a / sqrt(2.0 * a - 1.0);
                re -= w * v.x;
                im -= w * v.y;
            }
            q[x] = a + 1;
            {
                const zc v = t[it.plus[k] + at(np, row * ncp + comp_index(q[0], q[1], q[2]))];
                const double w = sqrt(2.0 * a + 1.0);
                re += w * v.x;
                im += w * v.y;
            }
            q[x] = a;
            if (phase) {
                // (B x e_x)_y for y = 0, 1, 2
                const double cr[3] = {x == 0 ? 0.0 : (x == 1 ? -B[2] : B[1]),
                                      x == 1 ? 0.0 : (x == 0 ? B[2] : -B[0]),
                                      x == 2 ? 0.0 : (x == 0 ? -B[1] : B[0])};
                for (int y = 0; y < 3; ++y) {
                    const double coef = sign * cr[y];
                    if (coef == 0.0) continue;
                    q[y] += 1;
                    const zc v = t[it.plus[k] + at(np, (it.krow[k] + row) * ncp + comp_index(q[0], q[1], q[2]))];
                    q[y] -= 1;
                    const double w = sqrt(2.0 * pw[y] + 1.0);
                    const double cvr = w * v.x + it.center[k][y] * base.x;
                    const double cvi = w * v.y + it.center[k][y] * base.y;
                    re -= coef * cvi;
                    im += coef * cvr;
                }
            }
            dst[(long long)(k * 3 + x) * it.size + e] = zc{re, im};
        }
    }
}

}  // namespace

hipError_t launch_deriv_assemble(const DerivItem* items, const int* tile_item, int ntiles, const zc* t, zc* out,
                                 const double B[3], bool phase, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    deriv_assemble_kernel<<<ntiles, 256, 0, s>>>(items, tile_item, t, out, B[0], B[1], B[2], phase ? 1 : 0);
    return hipGetLastError();
}

}  // namespace gh
