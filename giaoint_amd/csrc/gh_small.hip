// Lane-per-primitive-quartet ERI kernel for the smallest classes
// (la + lb + lc + ld <= 2): the same E.R.E formulation as the generic
// transform (gh_eri.hip), but one wavefront per work item whose lanes take
// the item's primitive quartets in turn, each lane building its own R0 table
// and one-dimensional Hermite tables in registers and accumulating the whole
// (small) Cartesian block; one wavefront reduction at the end. Every index is
// a compile-time constant of the class, so nothing is staged through LDS and
// no lane waits on another before the reduction.
//
// Reference: the contraction of src/eri.cpp:341-368 per component quartet,
// with the London phase folded into the Hermite coefficients as in the other
// kernels (Et[k] = sum_t E_t C(t,k) (-iq)^(t-k)).
#include "gh_hermite.h"
#include "gh_internal.h"
#include "gh_plan.h"

namespace gh {
namespace sm {

// Complex 1-D Hermite table of a pair on one axis, shell powers I, J: the
// field-free recurrence (raise the bra power I times with PA, then the ket
// power J times with PB; hermite.cpp:8-40), folded with the wave number q.
template <int I, int J>
__device__ inline void hermite_1d(double PA, double PB, double p, double q, zc (&Et)[I + J + 1]) {
    constexpr int T = I + J;
    double E[T + 1];
#pragma unroll
    for (int t = 0; t <= T; ++t) E[t] = t == 0 ? 1.0 : 0.0;
    const double i2p = 0.5 / p;
#pragma unroll
    for (int step = 0; step < T; ++step) {
        const double disp = step < I ? PA : PB;
        double nxt[T + 1];
#pragma unroll
        for (int t = 0; t <= T; ++t) {
            const double lower = t > 0 ? E[t - 1] : 0.0;
            const double same = t <= step ? E[t] : 0.0;
            const double upper = t + 1 <= step ? E[t + 1 <= T ? t + 1 : T] : 0.0;
            nxt[t] = t <= step + 1 ? i2p * lower + disp * same + (double)(t + 1) * upper : 0.0;
        }
#pragma unroll
        for (int t = 0; t <= T; ++t) E[t] = nxt[t];
    }
#pragma unroll
    for (int k = 0; k <= T; ++k) {
        double re = 0.0, im = 0.0, qp = 1.0;
#pragma unroll
        for (int m = 0; k + m <= T; ++m) {
            const double v = E[k + m] * binom(k + m, k) * qp;
            if ((m & 3) == 0) re += v;
            else if ((m & 3) == 1) im -= v;
            else if ((m & 3) == 2) re -= v;
            else im += v;
            qp *= q;
        }
        Et[k] = zc{re, im};
    }
}

// Hermite tables of a pair for every power combination: H[axis][i][j][k].
template <int LA, int LB>
struct PairTables {
    zc h[3][LA + 1][LB + 1][LA + LB + 1];
};

template <int I, int J, int LA, int LB>
__device__ inline void fill_pair(PairTables<LA, LB>& T, int axis, double PA, double PB, double p, double q) {
    zc tmp[I + J + 1];
    hermite_1d<I, J>(PA, PB, p, q, tmp);
#pragma unroll
    for (int k = 0; k <= LA + LB; ++k) T.h[axis][I][J][k] = k <= I + J ? tmp[k] : zc{0.0, 0.0};
}

template <int LA, int LB, int I = 0, int J = 0>
__device__ inline void fill_all(PairTables<LA, LB>& T, int axis, double PA, double PB, double p, double q) {
    if constexpr (I <= LA) {
        if constexpr (J <= LB) {
            fill_pair<I, J, LA, LB>(T, axis, PA, PB, p, q);
            fill_all<LA, LB, I, J + 1>(T, axis, PA, PB, p, q);
        } else {
            fill_all<LA, LB, I + 1, 0>(T, axis, PA, PB, p, q);
        }
    }
}

// R0 table of degree <= L at the complex displacement X from the seeds
// s_0..s_L (levels n = L..0), lexicographic order.
template <int L>
__device__ inline void r0_table(const zc (&X)[3], const zc (&s)[L + 1], zc (&R)[plan::nh(L)]) {
    constexpr int NH = plan::nh(L);
    zc cur[NH];
#pragma unroll
    for (int e = 0; e < NH; ++e) cur[e] = zc{0.0, 0.0};
    cur[0] = s[L];
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
        zc nxt[NH];
#pragma unroll
        for (int t = 0; t <= L; ++t)
#pragma unroll
            for (int u = 0; t + u <= L; ++u)
#pragma unroll
                for (int v = 0; t + u + v <= L; ++v) {
                    const int e = plan::lexidx(L, t, u, v);
                    zc val{0.0, 0.0};
                    if (t + u + v <= L - n) {
                        if (t > 0) {
                            val = zmul(X[0], cur[plan::lexidx(L, t - 1, u, v)]);
                            if (t > 1) {
                                const zc r2 = cur[plan::lexidx(L, t > 1 ? t - 2 : 0, u, v)];
                                val.x += (double)(t - 1) * r2.x;
                                val.y += (double)(t - 1) * r2.y;
                            }
                        } else if (u > 0) {
                            val = zmul(X[1], cur[plan::lexidx(L, 0, u - 1, v)]);
                            if (u > 1) {
                                const zc r2 = cur[plan::lexidx(L, 0, u > 1 ? u - 2 : 0, v)];
                                val.x += (double)(u - 1) * r2.x;
                                val.y += (double)(u - 1) * r2.y;
                            }
                        } else if (v > 0) {
                            val = zmul(X[2], cur[v - 1]);
                            if (v > 1) {
                                const zc r2 = cur[v > 1 ? v - 2 : 0];
                                val.x += (double)(v - 1) * r2.x;
                                val.y += (double)(v - 1) * r2.y;
                            }
                        } else {
                            val = s[n];
                        }
                    }
                    nxt[e] = val;
                }
#pragma unroll
        for (int e = 0; e < NH; ++e) cur[e] = nxt[e];
    }
#pragma unroll
    for (int e = 0; e < NH; ++e) R[e] = cur[e];
}

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void small_kernel(const WorkItem* __restrict__ items,
                                                  const DevShell* __restrict__ shells,
                                                  const double* __restrict__ exps, const double* __restrict__ coefs,
                                                  const double* __restrict__ norms, const double* __restrict__ kappa,
                                                  const zc* __restrict__ seeds, zc* __restrict__ out) {
    constexpr int L = LA + LB + LC + LD;
    constexpr int NCA = plan::nc(LA), NCB = plan::nc(LB), NCC = plan::nc(LC), NCD = plan::nc(LD);
    constexpr int NH = plan::nh(L);
    const WorkItem it = items[blockIdx.x];
    const int lane = threadIdx.x;
    const DevShell& A = shells[it.sh[0]];
    const DevShell& B = shells[it.sh[1]];
    const DevShell& C = shells[it.sh[2]];
    const DevShell& D = shells[it.sh[3]];
    const int Ka = A.nprim, Kb = B.nprim, Kc = C.nprim, Kd = D.nprim;
    const int Kcd = Kc * Kd, K = Ka * Kb * Kcd;
    double Ac[3], Bc[3], Cc[3], Dc[3], qab[3], qcd[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) {
        Ac[x] = A.center[x];
        Bc[x] = B.center[x];
        Cc[x] = C.center[x];
        Dc[x] = D.center[x];
        qab[x] = kappa[3 * it.sh[1] + x] - kappa[3 * it.sh[0] + x];
        qcd[x] = kappa[3 * it.sh[3] + x] - kappa[3 * it.sh[2] + x];
    }
    zc acc[NCA * NCB][NCC * NCD];
#pragma unroll
    for (int ab = 0; ab < NCA * NCB; ++ab)
#pragma unroll
        for (int cd = 0; cd < NCC * NCD; ++cd) acc[ab][cd] = zc{0.0, 0.0};

    for (int pq = lane; pq < K; pq += 64) {
        const int pab = pq / Kcd, pcd = pq - pab * Kcd;
        const int pa = pab / Kb, pb = pab - pa * Kb, pc = pcd / Kd, pd = pcd - pc * Kd;
        const double ea = exps[A.prim_off + pa], eb = exps[B.prim_off + pb];
        const double ec = exps[C.prim_off + pc], ed = exps[D.prim_off + pd];
        const zc* sp = seeds + it.seed_off + (long long)pq * (L + 4);
        zc X[3], s[L + 1];
#pragma unroll
        for (int x = 0; x < 3; ++x) X[x] = sp[x];
#pragma unroll
        for (int n = 0; n <= L; ++n) s[n] = sp[3 + n];
        // bra and ket Hermite tables (the same expressions as the generic kernel's builder)
        PairTables<LA, LB> Tb;
        PairTables<LC, LD> Tk;
        const double p = ea + eb, q = ec + ed;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            const double Pc = (ea * Ac[x] + eb * Bc[x]) / p;
            fill_all<LA, LB>(Tb, x, Pc - Ac[x], Pc - Bc[x], p, qab[x]);
            const double Qc = (ec * Cc[x] + ed * Dc[x]) / q;
            fill_all<LC, LD>(Tk, x, Qc - Cc[x], Qc - Dc[x], q, qcd[x]);
        }
        zc R[NH];
        r0_table<L>(X, s, R);
        // contraction coefficients x normalizations of the four primitives
        double fa[NCA], fb[NCB], fc[NCC], fd[NCD];
        const double ca0 = coefs[A.coef_off + it.x[0] * Ka + pa], cb0 = coefs[B.coef_off + it.x[1] * Kb + pb];
        const double cc0 = coefs[C.coef_off + it.x[2] * Kc + pc], cd0 = coefs[D.coef_off + it.x[3] * Kd + pd];
#pragma unroll
        for (int c = 0; c < NCA; ++c) fa[c] = ca0 * norms[A.norm_off + c * Ka + pa];
#pragma unroll
        for (int c = 0; c < NCB; ++c) fb[c] = cb0 * norms[B.norm_off + c * Kb + pb];
#pragma unroll
        for (int c = 0; c < NCC; ++c) fc[c] = cc0 * norms[C.norm_off + c * Kc + pc];
#pragma unroll
        for (int c = 0; c < NCD; ++c) fd[c] = cd0 * norms[D.norm_off + c * Kd + pd];
        // (ab|cd) = sum_{tuv} Et_ab sum_{tau phi chi} (-1)^(tau+phi+chi) Et_cd R0[t+tau, u+phi, v+chi]
#pragma unroll
        for (int cc = 0; cc < NCC; ++cc)
#pragma unroll
            for (int dc = 0; dc < NCD; ++dc) {
                const plan::P3 c3 = plan::cart(LC, cc), d3 = plan::cart(LD, dc);
                // W[tuv] for every bra Hermite index of degree <= LA + LB
                zc W[plan::nh(LA + LB)];
#pragma unroll
                for (int t = 0; t <= LA + LB; ++t)
#pragma unroll
                    for (int u = 0; t + u <= LA + LB; ++u)
#pragma unroll
                        for (int v = 0; t + u + v <= LA + LB; ++v) {
                            zc w{0.0, 0.0};
#pragma unroll
                            for (int ta = 0; ta <= c3.x + d3.x; ++ta)
#pragma unroll
                                for (int ph = 0; ph <= c3.y + d3.y; ++ph)
#pragma unroll
                                    for (int ch = 0; ch <= c3.z + d3.z; ++ch) {
                                        const zc e = zmul(zmul(Tk.h[0][c3.x][d3.x][ta], Tk.h[1][c3.y][d3.y][ph]),
                                                          Tk.h[2][c3.z][d3.z][ch]);
                                        const zc r = R[plan::lexidx(L, t + ta, u + ph, v + ch)];
                                        if ((ta + ph + ch) & 1) w = zsub(w, zmul(e, r));
                                        else w = zadd(w, zmul(e, r));
                                    }
                            W[plan::lexidx(LA + LB, t, u, v)] = w;
                        }
                const double fcd = fc[cc] * fd[dc];
#pragma unroll
                for (int ac = 0; ac < NCA; ++ac)
#pragma unroll
                    for (int bc = 0; bc < NCB; ++bc) {
                        const plan::P3 a3 = plan::cart(LA, ac), b3 = plan::cart(LB, bc);
                        zc v{0.0, 0.0};
#pragma unroll
                        for (int t = 0; t <= a3.x + b3.x; ++t)
#pragma unroll
                            for (int u = 0; u <= a3.y + b3.y; ++u)
#pragma unroll
                                for (int w3 = 0; w3 <= a3.z + b3.z; ++w3) {
                                    const zc e = zmul(zmul(Tb.h[0][a3.x][b3.x][t], Tb.h[1][a3.y][b3.y][u]),
                                                      Tb.h[2][a3.z][b3.z][w3]);
                                    v = zadd(v, zmul(e, W[plan::lexidx(LA + LB, t, u, w3)]));
                                }
                        const double f = fa[ac] * fb[bc] * fcd;
                        zc& a = acc[ac * NCB + bc][cc * NCD + dc];
                        a.x = fma(f, v.x, a.x);
                        a.y = fma(f, v.y, a.y);
                    }
            }
    }
    // sum over the wavefront's lanes (butterfly: every lane ends with the sums)
#pragma unroll
    for (int ab = 0; ab < NCA * NCB; ++ab)
#pragma unroll
        for (int cd = 0; cd < NCC * NCD; ++cd)
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                acc[ab][cd].x += __shfl_xor(acc[ab][cd].x, o);
                acc[ab][cd].y += __shfl_xor(acc[ab][cd].y, o);
            }
    // lane o writes output element o of the item's ket range
#pragma unroll
    for (int ab = 0; ab < NCA * NCB; ++ab)
#pragma unroll
        for (int cd = 0; cd < NCC * NCD; ++cd) {
            const int o = ab * NCC * NCD + cd;
            if (lane == o % 64 && cd >= it.cd0 && cd < it.cd1) {
                const long long ia = (long long)it.x[0] * NCA + ab / NCB, ib = (long long)it.x[1] * NCB + ab % NCB;
                const long long ic = (long long)it.x[2] * NCC + cd / NCD, id = (long long)it.x[3] * NCD + cd % NCD;
                out[it.out_base + ia * it.stride[0] + ib * it.stride[1] + ic * it.stride[2] + id * it.stride[3]] =
                    acc[ab][cd];
            }
        }
}

}  // namespace sm

bool small_class(int la, int lb, int lc, int ld) { return la + lb + lc + ld <= kSmallMaxL; }

hipError_t launch_small(int la, int lb, int lc, int ld, const WorkItem* items, int n, const DevShell* shells,
                        const double* exps, const double* coefs, const double* norms, const double* kappa,
                        const zc* seeds, zc* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
#define GH_SM(A, B, C, D)                                                                                        \
    if (la == A && lb == B && lc == C && ld == D) {                                                              \
        sm::small_kernel<A, B, C, D><<<dim3((unsigned)n), dim3(64), 0, s>>>(items, shells, exps, coefs, norms,   \
                                                                           kappa, seeds, out);                   \
        return hipGetLastError();                                                                                \
    }
    GH_SM(0, 0, 0, 0)
    GH_SM(1, 0, 0, 0)
    GH_SM(0, 1, 0, 0)
    GH_SM(0, 0, 1, 0)
    GH_SM(0, 0, 0, 1)
    GH_SM(1, 1, 0, 0)
    GH_SM(1, 0, 1, 0)
    GH_SM(1, 0, 0, 1)
    GH_SM(0, 1, 1, 0)
    GH_SM(0, 1, 0, 1)
    GH_SM(0, 0, 1, 1)
    GH_SM(2, 0, 0, 0)
    GH_SM(0, 2, 0, 0)
    GH_SM(0, 0, 2, 0)
    GH_SM(0, 0, 0, 2)
#undef GH_SM
    return hipErrorInvalidValue;
}

}  // namespace gh
