// Class-specialized (batched) E.R.E transform for gfx950.
//
// Same mathematics as transform_kernel (gh_eri.hip): per primitive quartet
// an R0 table at the complex displacement, the ket transform
// W[k][cd] += sum_j Et_cd[j] (-1)^|j| R0[k+j] accumulated over ket primitive
// pairs, then out[ab][cd] += sum_k Et_ab[k] W[k][cd] per bra primitive pair.
// What changes is the mapping onto the machine:
//
//  * pair kernel: the complex Hermite products Et_x Et_y Et_z of every
//    primitive pair are written once, in the compile-time order of a plan
//    (gh_plan.h), with the (-1)^|j| sign folded in for the ket side;
//  * transform: lane = bra Hermite index k; W[k][cd] for a chunk of ket
//    components lives in registers; the product for (j, cd) is identical in
//    every lane, so it is read with a wave-uniform (scalar) load and one R0
//    value fetched from LDS feeds every component whose box contains j;
//  * the four waves of a workgroup take different ket primitive pairs with
//    wave-level synchronization only; their W partials are summed through LDS
//    in a fixed order (bitwise reproducible), and the bra step splits the
//    bra components round-robin over the waves (lane = ket component);
//  * contraction coefficients scale the seeds when staged in LDS and the
//    primitive normalization is split into a radial factor (folded into the
//    seeds by the seed kernel) and a component factor (applied on output).
#include "gh_hermite.h"
#include "gh_internal.h"
#include "gh_plan.h"

#include <array>

namespace gh {
namespace fast {

constexpr int kW = 4;
constexpr int kNT = 64 * kW;

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int LA, int LB, int LC, int LD>
struct Cls {
    static constexpr int la = LA, lb = LB, lc = LC, ld = LD;
    static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    static constexpr int NCA = plan::nc(LA), NCB = plan::nc(LB), NCC = plan::nc(LC), NCDD = plan::nc(LD);
    static constexpr int NAB = NCA * NCB, NCD = NCC * NCDD;
    static constexpr int NHAB = plan::nh(LAB), NHCD = plan::nh(LCD), NHL = plan::nh(L);
    static constexpr int RK = (NHAB + 63) / 64;
    static constexpr int TILE_MAX = 40 / RK;
    static constexpr int NCHUNK = (NCD + TILE_MAX - 1) / TILE_MAX;
    using Ket = plan::Side<LC, LD, NCHUNK, 0>;
    using Bra = plan::Side<LA, LB, kW, 1>;
    static constexpr int TILE = Ket::CH;
    static constexpr int NOUT = (NAB + kW - 1) / kW;
    static constexpr int NTP = (LCD + 1) * (LCD + 2) / 2;  // (tau, phi) pairs of the ket space
    static constexpr int WAVE_ZC = 2 * NHL + L + 4;
    static constexpr int WSH_ZC = NHAB * TILE;
    static constexpr int LDS_BYTES = (kW * WAVE_ZC + WSH_ZC) * 16;
    static constexpr int tp(int t, int p) { return t * (LCD + 1) - t * (t - 1) / 2 + p; }
};

// R0 entries of degree <= L sorted by degree: t | u << 6 | v << 12 | lexidx << 18.
template <int L>
struct RList {
    int e[plan::nh(L)];
};
template <int L>
constexpr RList<L> make_rlist() {
    RList<L> r{};
    int k = 0;
    for (int d = 0; d <= L; ++d)
        for (int t = d; t >= 0; --t)
            for (int u = d - t; u >= 0; --u) {
                const int v = d - t - u;
                r.e[k++] = t | (u << 6) | (v << 12) | (plan::lexidx(L, t, u, v) << 18);
            }
    return r;
}
template <int L>
__constant__ RList<L> c_rlist = make_rlist<L>();

// ---------------------------------------------------------------------------
// Pair kernel: products of one primitive pair in plan order.
// ---------------------------------------------------------------------------
template <int L1, int L2, int NG, int MODE>
__constant__ typename plan::Side<L1, L2, NG, MODE>::Tables c_side = plan::Side<L1, L2, NG, MODE>::make();

template <int L1, int L2, int NG, int MODE, bool SIGN>
__global__ __launch_bounds__(64) void pair_kernel(const FastPairQuartet* __restrict__ pq, int nq, long long total,
                                                  int side, const DevShell* __restrict__ shells,
                                                  const double* __restrict__ exps, const double* __restrict__ kappa,
                                                  zc* __restrict__ table) {
    using S = plan::Side<L1, L2, NG, MODE>;
    constexpr int D = L1 + L2;
    constexpr int NT1D = 3 * (L1 + 1) * (L2 + 1);
    __shared__ zc et[NT1D * (D + 1)];
    const long long unit = blockIdx.x;
    if (unit >= total) return;
    int lo = 0, hi = nq - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pq[mid].unit_begin[side] <= unit) lo = mid;
        else hi = mid - 1;
    }
    const FastPairQuartet& Q = pq[lo];
    const int local = (int)(unit - Q.unit_begin[side]);
    const int s1 = Q.sh[2 * side], s2 = Q.sh[2 * side + 1];
    const DevShell& A = shells[s1];
    const DevShell& B = shells[s2];
    const int p1 = local / B.nprim, p2 = local % B.nprim;
    const double ea = exps[A.prim_off + p1], eb = exps[B.prim_off + p2];
    const double p = ea + eb;
    for (int e = threadIdx.x; e < NT1D; e += 64) {
        const int axis = e / ((L1 + 1) * (L2 + 1));
        const int i = (e / (L2 + 1)) % (L1 + 1), j = e % (L2 + 1);
        const double P = (ea * A.center[axis] + eb * B.center[axis]) / p;
        const double q = kappa[3 * s2 + axis] - kappa[3 * s1 + axis];
        build_phase_hermite(i, j, P - A.center[axis], P - B.center[axis], p, q, et + e * (D + 1));
    }
    __syncthreads();
    zc* out = table + Q.table_off[side] + (long long)local * S::NE;
    for (int e = threadIdx.x; e < S::NE; e += 64) {
        const int c = c_side<L1, L2, NG, MODE>.comp[e];
        const int hx = c_side<L1, L2, NG, MODE>.hidx[e];
        const plan::P3 h = plan::herm(D, hx);
        const int ca = c / plan::nc(L2), cb = c % plan::nc(L2);
        const plan::P3 a = plan::cart(L1, ca), b = plan::cart(L2, cb);
        const zc ex = et[((0 * (L1 + 1) + a.x) * (L2 + 1) + b.x) * (D + 1) + h.x];
        const zc ey = et[((1 * (L1 + 1) + a.y) * (L2 + 1) + b.y) * (D + 1) + h.y];
        const zc ez = et[((2 * (L1 + 1) + a.z) * (L2 + 1) + b.z) * (D + 1) + h.z];
        zc v = zmul(zmul(ex, ey), ez);
        if (SIGN && ((h.x + h.y + h.z) & 1)) v = zneg(v);
        out[e] = v;
    }
}

// ---------------------------------------------------------------------------
// Transform kernel
// ---------------------------------------------------------------------------
template <class C, int CI, int E, int END>
__device__ inline void ket_entries(zc (&W)[C::RK][C::TILE], const zc (&rv)[C::RK], const zc* __restrict__ ek) {
    if constexpr (E < END) {
        constexpr int cl = C::Ket::T.comp[E] - CI * C::TILE;
        const zc ev = ek[E];
#pragma unroll
        for (int r = 0; r < C::RK; ++r) zmac(W[r][cl], ev, rv[r]);
        ket_entries<C, CI, E + 1, END>(W, rv, ek);
    }
}

// W[k][cd] += sum_j E_cd[j] R0[k + j] for the lane's bra Hermite indices k.
template <class C, int CI, int H>
__device__ inline void ket_step(zc (&W)[C::RK][C::TILE], const zc* __restrict__ R0, const int (&kb)[C::RK][C::NTP],
                                const zc* __restrict__ ek) {
    if constexpr (H < C::NHCD) {
        constexpr plan::P3 j = plan::herm(C::LCD, H);
        constexpr int b = C::Ket::T.start[CI][H], e = C::Ket::T.start[CI][H + 1];
        if constexpr (e > b) {
            zc rv[C::RK];
#pragma unroll
            for (int r = 0; r < C::RK; ++r) rv[r] = R0[kb[r][C::tp(j.x, j.y)] + j.z];
            ket_entries<C, CI, b, e>(W, rv, ek);
        }
        ket_step<C, CI, H + 1>(W, R0, kb, ek);
    }
}

template <class C, int WV, int E, int END>
__device__ inline void bra_entries(zc (&OUT)[C::NOUT], const zc wv, const zc* __restrict__ eb) {
    if constexpr (E < END) {
        constexpr int slot = C::Bra::T.comp[E] / kW;
        zmac(OUT[slot], eb[E], wv);
        bra_entries<C, WV, E + 1, END>(OUT, wv, eb);
    }
}

// OUT[ab] += sum_k E_ab[k] W[k][cd] for the wave's bra components (lane = cd).
template <class C, int WV, int H>
__device__ inline void bra_step(zc (&OUT)[C::NOUT], const zc* __restrict__ Wsh, int cl, const zc* __restrict__ eb) {
    if constexpr (H < C::NHAB) {
        constexpr int b = C::Bra::T.start[WV][H], e = C::Bra::T.start[WV][H + 1];
        if constexpr (e > b) {
            const zc wv = Wsh[H * C::TILE + cl];
            bra_entries<C, WV, b, e>(OUT, wv, eb);
        }
        bra_step<C, WV, H + 1>(OUT, Wsh, cl, eb);
    }
}

// Component parts of the primitive normalizations.
template <class C>
struct Norms {
    std::array<double, C::NAB> ab{};
    std::array<double, C::NCD> cd{};
};
template <class C>
constexpr Norms<C> make_norms() {
    Norms<C> n{};
    for (int c = 0; c < C::NAB; ++c)
        n.ab[c] = plan::angular_norm(plan::cart(C::la, c / C::NCB)) * plan::angular_norm(plan::cart(C::lb, c % C::NCB));
    for (int c = 0; c < C::NCD; ++c)
        n.cd[c] = plan::angular_norm(plan::cart(C::lc, c / C::NCDD)) * plan::angular_norm(plan::cart(C::ld, c % C::NCDD));
    return n;
}
template <class C>
__constant__ Norms<C> c_norms = make_norms<C>();

template <class C, int WV, int AB>
__device__ inline void write_ab(const zc (&OUT)[C::NOUT], double gcd, long long base, const FastItem& it,
                                zc* __restrict__ out) {
    if constexpr (AB < C::NAB) {
        constexpr double g = make_norms<C>().ab[AB];
        constexpr int ca = AB / C::NCB, cb = AB % C::NCB;
        zc v = zscale(g * gcd, OUT[AB / kW]);
        if (it.conj) v.y = -v.y;
        const long long ia = (long long)it.x[0] * C::NCA + ca, ib = (long long)it.x[1] * C::NCB + cb;
        out[base + ia * it.stride[0] + ib * it.stride[1]] = v;
        write_ab<C, WV, AB + kW>(OUT, gcd, base, it, out);
    }
}

template <class C, int WV>
__device__ inline void write_out(const zc (&OUT)[C::NOUT], int lane, int c0, int nthis, const FastItem& it,
                                 zc* __restrict__ out) {
    if (lane >= nthis) return;
    const int cd = c0 + lane;
    const int cc = cd / C::NCDD, dc = cd % C::NCDD;
    const double gcd = c_norms<C>.cd[cd];
    const long long ic = (long long)it.x[2] * C::NCC + cc, id = (long long)it.x[3] * C::NCDD + dc;
    const long long base = it.out_base + ic * it.stride[2] + id * it.stride[3];
    write_ab<C, WV, WV>(OUT, gcd, base, it, out);
}

template <class C, int CI>
__device__ inline void run_chunk(const FastItem& it, const DevShell& A, const DevShell& B, const DevShell& Cs,
                                 const DevShell& D, const double* __restrict__ coefs, const zc* __restrict__ seeds,
                                 const zc* __restrict__ etab, zc* __restrict__ out, zc* wave_lds, zc* Wsh, int wave,
                                 int lane, const int (&kb)[C::RK][C::NTP]) {
    constexpr int c0 = CI * C::TILE;
    constexpr int nthis = (C::NCD - c0) < C::TILE ? (C::NCD - c0) : C::TILE;
    const int Kb = B.nprim, Kd = D.nprim;
    const int Kab = A.nprim * Kb, Kcd = Cs.nprim * Kd;
    zc OUT[C::NOUT];
#pragma unroll
    for (int i = 0; i < C::NOUT; ++i) OUT[i] = zc{0.0, 0.0};
    zc* bufs[2] = {wave_lds, wave_lds + C::NHL};  // level m of R0 lives in bufs[m & 1]
    zc* sseed = wave_lds + 2 * C::NHL;
    for (int pab = 0; pab < Kab; ++pab) {
        const int pa = pab / Kb, pb = pab % Kb;
        const double cab = coefs[A.coef_off + it.x[0] * A.nprim + pa] * coefs[B.coef_off + it.x[1] * B.nprim + pb];
        zc W[C::RK][C::TILE];
#pragma unroll
        for (int r = 0; r < C::RK; ++r)
#pragma unroll
            for (int t = 0; t < C::TILE; ++t) W[r][t] = zc{0.0, 0.0};
        for (int pcd = wave; pcd < Kcd; pcd += kW) {
            const int pc = pcd / Kd, pd = pcd % Kd;
            const double sc = cab * coefs[Cs.coef_off + it.x[2] * Cs.nprim + pc] *
                              coefs[D.coef_off + it.x[3] * D.nprim + pd];
            const zc* sp = seeds + it.seed_off + (long long)(pab * Kcd + pcd) * (C::L + 4);
            if (lane < C::L + 4) {
                const zc v = sp[lane];
                sseed[lane] = lane < 3 ? v : zscale(sc, v);
            }
            wave_sync();
            const zc X0 = sseed[0], X1 = sseed[1], X2 = sseed[2];
            if (lane == 0) bufs[C::L & 1][0] = sseed[3 + C::L];
            // level n from level n + 1 (R0^n_{t+1} = t R0^{n+1}_{t-1} + X R0^{n+1}_t)
#pragma unroll
            for (int n = C::L - 1; n >= 0; --n) {
                wave_sync();
                const zc* cur = bufs[(n + 1) & 1];
                zc* nxt = bufs[n & 1];
                const int cnt = plan::nh(C::L - n);
#pragma unroll
                for (int base = 0; base < C::NHL; base += 64) {
                    const int e = base + lane;
                    if (base < cnt && e < cnt) {
                        const int pk = c_rlist<C::L>.e[e];
                        const int t = pk & 63, u = (pk >> 6) & 63, v = (pk >> 12) & 63, lex = pk >> 18;
                        zc val;
                        if (t > 0) {
                            val = zmul(X0, cur[plan::lexoff(C::L, t - 1, u) + v]);
                            if (t > 1) {
                                const zc r2 = cur[plan::lexoff(C::L, t - 2, u) + v];
                                val.x += (double)(t - 1) * r2.x;
                                val.y += (double)(t - 1) * r2.y;
                            }
                        } else if (u > 0) {
                            val = zmul(X1, cur[plan::lexoff(C::L, 0, u - 1) + v]);
                            if (u > 1) {
                                const zc r2 = cur[plan::lexoff(C::L, 0, u - 2) + v];
                                val.x += (double)(u - 1) * r2.x;
                                val.y += (double)(u - 1) * r2.y;
                            }
                        } else if (v > 0) {
                            val = zmul(X2, cur[v - 1]);
                            if (v > 1) {
                                const zc r2 = cur[v - 2];
                                val.x += (double)(v - 1) * r2.x;
                                val.y += (double)(v - 1) * r2.y;
                            }
                        } else {
                            val = sseed[3 + n];
                        }
                        nxt[lex] = val;
                    }
                }
            }
            wave_sync();
            const zc* ek = etab + it.eket_off + (long long)pcd * C::Ket::NE;
            ket_step<C, CI, 0>(W, bufs[0], kb, ek);
            wave_sync();
        }
        __syncthreads();
        for (int w = 0; w < kW; ++w) {
            if (wave == w) {
#pragma unroll
                for (int r = 0; r < C::RK; ++r) {
                    const int k = lane + 64 * r;
                    if (k < C::NHAB) {
#pragma unroll
                        for (int t = 0; t < C::TILE; ++t) {
                            zc* dst = Wsh + k * C::TILE + t;
                            if (w == 0) *dst = W[r][t];
                            else *dst = zadd(*dst, W[r][t]);
                        }
                    }
                }
            }
            __syncthreads();
        }
        const zc* eb = etab + it.ebra_off + (long long)pab * C::Bra::NE;
        const int cl = lane < C::TILE ? lane : C::TILE - 1;
        switch (wave) {
            case 0: bra_step<C, 0, 0>(OUT, Wsh, cl, eb); break;
            case 1: bra_step<C, 1, 0>(OUT, Wsh, cl, eb); break;
            case 2: bra_step<C, 2, 0>(OUT, Wsh, cl, eb); break;
            default: bra_step<C, 3, 0>(OUT, Wsh, cl, eb); break;
        }
        __syncthreads();
    }
    switch (wave) {
        case 0: write_out<C, 0>(OUT, lane, c0, nthis, it, out); break;
        case 1: write_out<C, 1>(OUT, lane, c0, nthis, it, out); break;
        case 2: write_out<C, 2>(OUT, lane, c0, nthis, it, out); break;
        default: write_out<C, 3>(OUT, lane, c0, nthis, it, out); break;
    }
}

template <class C, int CI>
struct ChunkLoop {
    __device__ static inline void run(const FastItem& it, const DevShell& A, const DevShell& B, const DevShell& Cs,
                                      const DevShell& D, const double* coefs, const zc* seeds, const zc* etab,
                                      zc* out, zc* wave_lds, zc* Wsh, int wave, int lane,
                                      const int (&kb)[C::RK][C::NTP]) {
        run_chunk<C, CI>(it, A, B, Cs, D, coefs, seeds, etab, out, wave_lds, Wsh, wave, lane, kb);
        if constexpr (CI + 1 < C::NCHUNK) {
            __syncthreads();
            ChunkLoop<C, CI + 1>::run(it, A, B, Cs, D, coefs, seeds, etab, out, wave_lds, Wsh, wave, lane, kb);
        }
    }
};

template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(kNT) void fast_transform_kernel(const FastItem* __restrict__ items,
                                                             const DevShell* __restrict__ shells,
                                                             const double* __restrict__ coefs,
                                                             const zc* __restrict__ seeds,
                                                             const zc* __restrict__ etab, zc* __restrict__ out) {
    using C = Cls<LA, LB, LC, LD>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    zc* lds = reinterpret_cast<zc*>(smem_raw);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int lane = threadIdx.x % 64;
    zc* wave_lds = lds + wave * C::WAVE_ZC;
    zc* Wsh = lds + kW * C::WAVE_ZC;
    const FastItem it = items[blockIdx.x];
    const DevShell A = shells[it.sh[0]];
    const DevShell B = shells[it.sh[1]];
    const DevShell Cs = shells[it.sh[2]];
    const DevShell D = shells[it.sh[3]];
    // per-lane R0 offsets: kb[r][(tau,phi)] = index(k_t + tau, k_u + phi, k_v)
    int kb[C::RK][C::NTP];
#pragma unroll
    for (int r = 0; r < C::RK; ++r) {
        int k = lane + 64 * r;
        if (k >= C::NHAB) k = 0;
        // decode lexicographic Hermite index k of degree <= LAB
        int kt = 0, ku = 0, kv = 0, rem = k;
        for (int t = 0; t <= C::LAB; ++t) {
            const int cnt_t = plan::nh(C::LAB - t) - plan::nh(C::LAB - t - 1);
            if (rem < cnt_t) {
                kt = t;
                for (int u = 0; u <= C::LAB - t; ++u) {
                    const int cnt_u = C::LAB - t - u + 1;
                    if (rem < cnt_u) {
                        ku = u;
                        kv = rem;
                        break;
                    }
                    rem -= cnt_u;
                }
                break;
            }
            rem -= cnt_t;
        }
#pragma unroll
        for (int t = 0; t <= C::LCD; ++t)
#pragma unroll
            for (int p = 0; p + t <= C::LCD; ++p) kb[r][C::tp(t, p)] = plan::lexoff(C::L, kt + t, ku + p) + kv;
    }
    ChunkLoop<C, 0>::run(it, A, B, Cs, D, coefs, seeds, etab, out, wave_lds, Wsh, wave, lane, kb);
}

// ---------------------------------------------------------------------------
// Launch table
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
hipError_t launch_class(const FastPairQuartet* pq, int nq, long long bra_units, long long ket_units,
                        const FastItem* items, int nitems, const DevShell* shells, const double* exps,
                        const double* coefs, const double* kappa, const zc* seeds, zc* etab, zc* out,
                        hipStream_t s) {
    using C = Cls<LA, LB, LC, LD>;
    if (bra_units)
        pair_kernel<LA, LB, kW, 1, false><<<dim3((unsigned)bra_units), dim3(64), 0, s>>>(pq, nq, bra_units, 0, shells,
                                                                                         exps, kappa, etab);
    if (ket_units)
        pair_kernel<LC, LD, C::NCHUNK, 0, true><<<dim3((unsigned)ket_units), dim3(64), 0, s>>>(pq, nq, ket_units, 1,
                                                                                               shells, exps, kappa,
                                                                                               etab);
    if (nitems) {
        static bool raised = false;
        if (!raised && C::LDS_BYTES > 64 * 1024) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&fast_transform_kernel<LA, LB, LC, LD>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, C::LDS_BYTES);
            if (e != hipSuccess) return e;
            raised = true;
        }
        fast_transform_kernel<LA, LB, LC, LD><<<dim3((unsigned)nitems), dim3(kNT), C::LDS_BYTES, s>>>(
            items, shells, coefs, seeds, etab, out);
    }
    return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
FastClassInfo info() {
    using C = Cls<LA, LB, LC, LD>;
    return FastClassInfo{true, C::Bra::NE, C::Ket::NE, C::LDS_BYTES, &launch_class<LA, LB, LC, LD>};
}

}  // namespace fast

// Classes with a specialized kernel. The orientation chosen by the host puts
// the pair with fewer components on the bra side.
FastClassInfo fast_class_info(int la, int lb, int lc, int ld) {
#define GH_FAST(a, b, c, d) \
    if (la == a && lb == b && lc == c && ld == d) return fast::info<a, b, c, d>();
    GH_FAST(2, 2, 2, 2)
    GH_FAST(2, 2, 3, 3)
    GH_FAST(1, 1, 1, 1)
    GH_FAST(1, 0, 1, 0)
    GH_FAST(0, 0, 0, 0)
#undef GH_FAST
    return FastClassInfo{false, 0, 0, 0, nullptr};
}

}  // namespace gh
