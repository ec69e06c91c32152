// Class-specialized (batched) E.R.E transform for gfx950: kernel templates.
// Each gh_fast_*.hip file instantiates the classes it owns.
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
//  * R0 kernel: one wavefront per primitive quartet runs the recurrence level
//    by level, a lane per (u, v) column of the table in registers, and stores
//    the final table (nh(L) entries) to HBM;
//  * transform: a workgroup of 4..8 wavefronts takes one work item and its
//    lanes are the rows (bra primitive pair, bra Hermite index k), so a dd bra
//    side with three-term contractions (9 x 35 rows) fills 5 wavefronts.
//    W[row][cd] for a chunk of ket components lives in registers. Per ket
//    primitive pair each lane gathers R0[k + j] with vector loads; the
//    products Et_cd[j], identical for every lane, come 16 plan entries at a
//    time: each lane of a 16-lane row loads one entry (one 16-byte load) and
//    the FMAs read entry n from lane n of the row by DPP row broadcast, so the
//    inner loop is FMAs with no separate operand traffic, software-pipelined
//    a few batches ahead (Pipe);
//  * bra step: the bra primitive pairs go one at a time through a
//    double-buffered LDS tile W[k][cd]; waves 0..3 each take a quarter of the
//    bra components (round robin), lane = ket component, products again by
//    row broadcast;
//  * contraction coefficients: ket pair coefficients scale the gathered R0
//    values, bra pair coefficients scale W on its way into LDS. The primitive
//    normalization is split into a radial factor (folded into the seeds by the
//    seed kernel) and a component factor (applied on output).
#pragma once

#include "gh_hermite.h"
#include "gh_internal.h"
#include "gh_plan.h"

#include <array>

namespace gh {
namespace fast {

constexpr int kG = kFastBraGroups;
constexpr int kMaxWaves = kFastMaxWaves;
constexpr int kMaxThreads = 64 * kMaxWaves;
constexpr int kTileMax = 40;                 // ket components per lane held in registers

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Returns v unchanged but opaque to the optimizer: values derived from it are
// computed where they are used instead of being hoisted out of loops into
// long-lived registers.
__device__ inline int opaque_v(int v) {
    asm volatile("" : "+v"(v));
    return v;
}
__device__ inline int opaque_s(int v) {
    asm volatile("" : "+s"(v));
    return v;
}
__device__ inline double opaque_d(double v) {
    asm volatile("" : "+v"(v));
    return v;
}

template <int LA, int LB, int LC, int LD>
struct Cls {
    static constexpr int la = LA, lb = LB, lc = LC, ld = LD;
    static constexpr int LAB = LA + LB, LCD = LC + LD, L = LAB + LCD;
    static constexpr int NCA = plan::nc(LA), NCB = plan::nc(LB), NCC = plan::nc(LC), NCDD = plan::nc(LD);
    static constexpr int NAB = NCA * NCB, NCD = NCC * NCDD;
    static constexpr int NHAB = plan::nh(LAB), NHCD = plan::nh(LCD), NHL = plan::nh(L);
    static constexpr int NCHUNK = (NCD + kTileMax - 1) / kTileMax;
    using Ket = plan::Side<LC, LD, NCHUNK, 0>;
    using Bra = plan::Side<LA, LB, kG, 1>;
    static constexpr int TILE = Ket::CH;
    static constexpr int WSTRIDE = TILE | 1;  // odd row stride (in complex) of the LDS tile
    static constexpr int WBUF = NHAB * WSTRIDE;
    static constexpr int NOUT = (NAB + kG - 1) / kG;
    static constexpr int NTP = (LCD + 1) * (LCD + 2) / 2;  // (tau, phi) pairs of the ket space
    static constexpr int tp(int t, int p) { return t * (LCD + 1) - t * (t - 1) / 2 + p; }
    static_assert(NHAB <= kMaxThreads, "bra Hermite rows exceed one workgroup");
    static_assert(TILE <= 64, "ket tile exceeds one wavefront");
};


// ---------------------------------------------------------------------------
// Pair kernel: products of one primitive pair in plan order.
// ---------------------------------------------------------------------------
// Per padded slot of a side's plan, the three 1-D Hermite table indices of
// its product (axis x, y, z: ((axis (L1+1) + a) (L2+1) + b) (D+1) + h) and
// the sign, packed as ix | iy << 10 | iz << 20 | neg << 30; -1 for padding.
template <int L1, int L2, int NG, int MODE, bool SIGN>
struct PairDesc {
    int d[plan::Side<L1, L2, NG, MODE>::NEP];
};
template <int L1, int L2, int NG, int MODE, bool SIGN>
constexpr PairDesc<L1, L2, NG, MODE, SIGN> make_pair_desc() {
    using S = plan::Side<L1, L2, NG, MODE>;
    constexpr int D = L1 + L2;
    static_assert(3 * (L1 + 1) * (L2 + 1) * (D + 1) <= 1024, "1-D table index exceeds 10 bits");
    constexpr auto T = S::make();
    PairDesc<L1, L2, NG, MODE, SIGN> r{};
    for (int slot = 0; slot < S::NEP; ++slot) {
        const int e = T.ent[slot];
        if (e == 0xffff) {
            r.d[slot] = -1;
            continue;
        }
        const int c = T.comp[e];
        const plan::P3 h = plan::herm(D, T.hidx[e]);
        const plan::P3 a = plan::cart(L1, c / plan::nc(L2)), b = plan::cart(L2, c % plan::nc(L2));
        const int ix = ((0 * (L1 + 1) + a.x) * (L2 + 1) + b.x) * (D + 1) + h.x;
        const int iy = ((1 * (L1 + 1) + a.y) * (L2 + 1) + b.y) * (D + 1) + h.y;
        const int iz = ((2 * (L1 + 1) + a.z) * (L2 + 1) + b.z) * (D + 1) + h.z;
        const int neg = SIGN && ((h.x + h.y + h.z) & 1) ? 1 : 0;
        r.d[slot] = ix | (iy << 10) | (iz << 20) | (neg << 30);
    }
    return r;
}
template <int L1, int L2, int NG, int MODE, bool SIGN>
__constant__ PairDesc<L1, L2, NG, MODE, SIGN> c_pair_desc = make_pair_desc<L1, L2, NG, MODE, SIGN>();

// One workgroup of four wavefronts per shell quartet; each wavefront takes
// the quartet's primitive pairs on one side in turn (side 0: bra pair (A, B),
// side 1: ket pair (C, D)), builds the pair's 1-D tables in its own LDS
// region and writes the products in plan order. A lane's slot descriptors
// are read once and serve every pair.
template <int L1, int L2, int NG, int MODE, bool SIGN>
__global__ __launch_bounds__(256) void pair_kernel(const FastPairQuartet* __restrict__ pq, int side,
                                                   const DevShell* __restrict__ shells,
                                                   const double* __restrict__ exps, const double* __restrict__ kappa,
                                                   zc* __restrict__ table) {
    using S = plan::Side<L1, L2, NG, MODE>;
    constexpr int D = L1 + L2;
    constexpr int NT1D = 3 * (L1 + 1) * (L2 + 1);
    constexpr int NS = (S::NEP + 63) / 64;
    __shared__ zc et_all[4][NT1D * (D + 1)];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
    zc* et = et_all[wave];
    int desc[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i)
        desc[i] = 64 * i + lane < S::NEP ? c_pair_desc<L1, L2, NG, MODE, SIGN>.d[64 * i + lane] : -2;
    const FastPairQuartet& Q = pq[blockIdx.x];
    const int s1 = Q.sh[2 * side], s2 = Q.sh[2 * side + 1];
    const DevShell& A = shells[s1];
    const DevShell& B = shells[s2];
    const int npairs = A.nprim * B.nprim;
    for (int local = wave; local < npairs; local += 4) {
        const int p1 = local / B.nprim, p2 = local % B.nprim;
        const double ea = exps[A.prim_off + p1], eb = exps[B.prim_off + p2];
        const double p = ea + eb;
        wave_sync();  // the previous pair's 1-D tables are no longer read
        for (int e = lane; e < NT1D; e += 64) {
            const int axis = e / ((L1 + 1) * (L2 + 1));
            const int i = (e / (L2 + 1)) % (L1 + 1), j = e % (L2 + 1);
            const double P = (ea * A.center[axis] + eb * B.center[axis]) / p;
            const double q = kappa[3 * s2 + axis] - kappa[3 * s1 + axis];
            build_phase_hermite(i, j, P - A.center[axis], P - B.center[axis], p, q, et + e * (D + 1));
        }
        wave_sync();
        zc* out = table + Q.table_off[side] + (long long)local * S::NEP;
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            const int d = desc[i];
            if (d == -2) continue;
            zc v = zc{0.0, 0.0};
            if (d >= 0) {
                v = zmul(zmul(et[d & 1023], et[(d >> 10) & 1023]), et[(d >> 20) & 1023]);
                if (d >> 30) v = zneg(v);
            }
            out[64 * i + lane] = v;
        }
    }
}

// ---------------------------------------------------------------------------
// R0 kernel: a wavefront builds the tables of G = 64 / ((L+1)(L+2)/2)
// primitive quartets at once, one per group of lanes. Seeds X (3) and
// s_0..s_L come from the seed kernel; levels n = L-1 .. 0 of
//   R^n_{t,u,v} = X0 R^{n+1}_{t-1,u,v} + (t-1) R^{n+1}_{t-2,u,v}    (t > 0)
//   R^n_{0,u,v} = X1 R^{n+1}_{0,u-1,v} + (u-1) R^{n+1}_{0,u-2,v}    (u > 0)
//   R^n_{0,0,v} = X2 R^{n+1}_{0,0,v-1} + (v-1) R^{n+1}_{0,0,v-2}    (v > 0)
//   R^n_{0,0,0} = s_n
// Lane c < (L+1)(L+2)/2 of a group owns the column (u, v) of the table and keeps its
// entries t = 0 .. L-n-u-v of the current and the previous level in
// registers: the t recurrence stays in the lane, and only the t = 0 entries
// of the two neighbouring columns cross lanes (shuffles), twice per level.
// Level 0 is stored in lexicographic order.
// ---------------------------------------------------------------------------
template <int L>
__global__ __launch_bounds__(256) void r0_kernel(const FastPairQuartet* __restrict__ pq,
                                                 const zc* __restrict__ seeds, zc* __restrict__ r0) {
    constexpr int NHL = plan::nh(L);
    constexpr int NCOL = (L + 1) * (L + 2) / 2;
    static_assert(NCOL <= 64, "one column per lane (r0_lds_kernel above that)");
    constexpr int G = 64 / NCOL;  // tables per wavefront, one per group of NCOL lanes
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
    const FastPairQuartet& Q = pq[blockIdx.x];
    const int units = Q.kab * Q.kcd;
    const int grp = lane / NCOL, col = lane - grp * NCOL;
    const bool own = grp < G;
    // the lane's column (u, v) and the lanes holding the columns its t = 0
    // entry is built from: (u-1, v), (u-2, v) or (0, v-1), (0, v-2), in its group
    int u = 0, v = 0;
    {
        int c = own ? col : 0, uu = 0;
        while (c > L - uu) {
            c -= L - uu + 1;
            ++uu;
        }
        u = uu;
        v = c;
    }
    auto column = [](int cu, int cv) { return cu * (L + 1) - cu * (cu - 1) / 2 + cv; };
    const int g0 = own ? grp * NCOL : 0;
    const int src1 = g0 + (u > 0 ? column(u - 1, v) : (v > 0 ? column(0, v - 1) : 0));
    const int src2 = g0 + (u > 1 ? column(u - 2, v) : (u == 0 && v > 1 ? column(0, v - 2) : 0));
    const double k2 = u > 0 ? (double)(u - 1) : (double)(v - 1);
    const bool origin = u == 0 && v == 0;
    // the wavefronts of the workgroup take the quartet's primitive quartets in
    // turn, G at a time (the loop is wave-uniform: every lane joins the shuffles)
    for (int base0 = wave * G; base0 < units; base0 += 4 * G) {
        const int local = base0 + (own ? grp : 0);
        const bool active = own && local < units;
        // the seeds this lane uses: X0, X1 or X2, and s_0..s_L in the origin
        // lane; with one table per wavefront the address is wave-uniform and
        // every seed is read into scalar registers
        const zc* sp = seeds + Q.seed_off + (long long)(G == 1 || !active ? base0 : local) * (L + 4);
        zc X0, Xb, sn[L + 1];
        if constexpr (G == 1) {
            zc sd[L + 4];
#pragma unroll
            for (int i = 0; i < L + 4; ++i) sd[i] = sp[i];
            X0 = sd[0];
            Xb = u > 0 ? sd[1] : sd[2];
#pragma unroll
            for (int i = 0; i <= L; ++i) sn[i] = sd[3 + i];
        } else {
            X0 = sp[0];
            Xb = u > 0 ? sp[1] : sp[2];
#pragma unroll
            for (int i = 0; i <= L; ++i) sn[i] = origin ? sp[3 + i] : zc{0.0, 0.0};
        }
        // the column at the current level, updated in place from high t down:
        // entries above the column's last (L - n - u - v) are never read by
        // entries that exist
        zc cur[L + 1];
#pragma unroll
        for (int t = 0; t <= L; ++t) cur[t] = zc{0.0, 0.0};
        if (origin) cur[0] = sn[L];
#pragma unroll
        for (int n = L - 1; n >= 0; --n) {
            const zc r1 = zc{__shfl(cur[0].x, src1), __shfl(cur[0].y, src1)};
            const zc r2 = zc{__shfl(cur[0].x, src2), __shfl(cur[0].y, src2)};
#pragma unroll
            for (int t = L - n; t >= 1; --t) {
                zc val = zmul(X0, cur[t - 1]);
                if (t > 1) {
                    val.x += (double)(t - 1) * cur[t - 2].x;
                    val.y += (double)(t - 1) * cur[t - 2].y;
                }
                cur[t] = val;
            }
            zc base = zmul(Xb, r1);
            if (k2 > 0.0) {
                base.x += k2 * r2.x;
                base.y += k2 * r2.y;
            }
            cur[0] = origin ? sn[n] : base;
        }
        if (active) {
            zc* dst = r0 + Q.r0_off + (long long)local * NHL;
            const int tmax = L - u - v;
#pragma unroll
            for (int t = 0; t <= L; ++t)
                if (t <= tmax) dst[plan::lexoff(L, t, u) + v] = cur[t];
        }
    }
}

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

// Tables with more than 64 columns (L >= 10): one wavefront per table runs
// the levels in LDS, each lane taking entries of the degree-sorted list.
// One R0 table: the lane's seed value (lane < L + 4: X then s_0..s_L) ->
// dst (nh(L) entries). pk holds the lane's entries e = 64 i + lane of the
// degree-sorted list (the same at every level: level n uses its first
// nh(L - n) entries), read from constant memory once per kernel.
template <int L>
__device__ inline void r0_table(const zc seed, zc* __restrict__ dst, zc* const (&bufs)[2], zc* sseed, int lane,
                                const int (&pk)[(plan::nh(L) + 63) / 64]) {
    constexpr int NHL = plan::nh(L);
    wave_sync();  // the previous table's LDS values are no longer read
    if (lane < L + 4) sseed[lane] = seed;
    wave_sync();
    const zc X0 = sseed[0], X1 = sseed[1], X2 = sseed[2];
    if (lane == 0) bufs[L & 1][0] = sseed[3 + L];
#pragma unroll
    for (int n = L - 1; n >= 0; --n) {
        wave_sync();
        const zc* cur = bufs[(n + 1) & 1];
        zc* nxt = bufs[n & 1];
        const int cnt = plan::nh(L - n);
#pragma unroll
        for (int base = 0; base < NHL; base += 64) {
            const int e = base + lane;
            if (base < cnt && e < cnt) {
                const int p = pk[base / 64];
                const int t = p & 63, u = (p >> 6) & 63, v = (p >> 12) & 63, lex = p >> 18;
                zc val;
                if (t > 0) {
                    val = zmul(X0, cur[plan::lexoff(L, t - 1, u) + v]);
                    if (t > 1) {
                        const zc r2 = cur[plan::lexoff(L, t - 2, u) + v];
                        val.x += (double)(t - 1) * r2.x;
                        val.y += (double)(t - 1) * r2.y;
                    }
                } else if (u > 0) {
                    val = zmul(X1, cur[plan::lexoff(L, 0, u - 1) + v]);
                    if (u > 1) {
                        const zc r2 = cur[plan::lexoff(L, 0, u - 2) + v];
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
    for (int e = lane; e < NHL; e += 64) dst[e] = bufs[0][e];
}

template <int L>
__global__ __launch_bounds__(256) void r0_lds_kernel(const FastPairQuartet* __restrict__ pq,
                                                 const zc* __restrict__ seeds, zc* __restrict__ r0) {
    constexpr int NHL = plan::nh(L);
    constexpr int NIT = (NHL + 63) / 64;
    constexpr int WZ = 2 * NHL + L + 4;
    __shared__ zc lds[4 * WZ];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
    const FastPairQuartet& Q = pq[blockIdx.x];
    const int units = Q.kab * Q.kcd;
    zc* wl = lds + wave * WZ;
    zc* bufs[2] = {wl, wl + NHL};
    zc* sseed = wl + 2 * NHL;
    int pk[NIT];
#pragma unroll
    for (int i = 0; i < NIT; ++i) pk[i] = 64 * i + lane < NHL ? c_rlist<L>.e[64 * i + lane] : 0;
    const zc* sbase = seeds + Q.seed_off + (lane < L + 4 ? lane : 0);
    // the four wavefronts of the workgroup take the quartet's primitive
    // quartets in turn; each prefetches its next table's seeds
    zc next = wave < units ? sbase[(long long)wave * (L + 4)] : zc{0.0, 0.0};
    for (int local = wave; local < units; local += 4) {
        const zc seed = next;
        if (local + 4 < units) next = sbase[(long long)(local + 4) * (L + 4)];
        r0_table<L>(seed, r0 + Q.r0_off + (long long)local * NHL, bufs, sseed, lane, pk);
    }
}


// ---------------------------------------------------------------------------
// Transform kernel
// ---------------------------------------------------------------------------
// acc += e(n) * r for complex values, where e(n) is the value that lane n of
// the lane's 16-lane row holds (DPP row_newbcast): every row holds the same
// batch of 16 plan entries, so e(n) is plan entry n of the batch in every lane.
// src0 of each FMA is the broadcast operand; the leading s_nop provides the
// wait states a DPP read needs after a VALU write of its source.
template <int N>
__device__ inline void zmac_bcast(zc& acc, const zc& e, const zc& r) {
    asm("s_nop 1\n\t"
        "v_fmac_f64_dpp %0, %2, %4 row_newbcast:%c6 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %1, %2, %5 row_newbcast:%c6 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %0, -%3, %5 row_newbcast:%c6 row_mask:0xf bank_mask:0xf\n\t"
        "v_fmac_f64_dpp %1, %3, %4 row_newbcast:%c6 row_mask:0xf bank_mask:0xf"
        : "+v"(acc.x), "+v"(acc.y)
        : "v"(e.x), "v"(e.y), "v"(r.x), "v"(r.y), "n"(N));
}

// Software-pipelined walk over the entries of plan group G, 16 entries (one
// load per lane) per batch. Batch B + D is loaded while batch B computes:
// its products and the per-lane values val[h] of the Hermite indices that
// batch B + D uses first (Src::load<h>), which are finished (Src::finish)
// when first used. Scheduling barriers keep the compiler from hoisting loads
// further, which bounds the registers in flight.
template <class P, int G, int D, class Slot>
struct Pipe {
    static constexpr int S = P::T.start[G][0], END = P::T.start[G][P::NH];
    static constexpr int NB = (END - S + 15) / 16;
    // The Hermite indices whose first entry lies in batch b, ascending:
    // h[start[b] .. start[b + 1]) (tabulated once per group).
    struct FirstUse {
        int start[NB + 1];
        int h[P::NH];
    };
    static constexpr FirstUse make_first_use() {
        FirstUse f{};
        int fb[P::NH];
        for (int h = 0; h < P::NH; ++h) fb[h] = -1;
        for (int e = S; e < END; ++e)
            if (fb[P::T.hidx[e]] < 0) fb[P::T.hidx[e]] = (e - S) / 16;
        int n = 0;
        for (int b = 0; b < NB; ++b) {
            f.start[b] = n;
            for (int h = 0; h < P::NH; ++h)
                if (fb[h] == b) f.h[n++] = h;
        }
        f.start[NB] = n;
        return f;
    }
    static constexpr FirstUse FU = make_first_use();

    template <int B, int I, class Src, int NV>
    __device__ static inline void load_vals(zc (&val)[NV], const Src& src) {
        if constexpr (FU.start[B] + I < FU.start[B + 1]) {
            constexpr int H = FU.h[FU.start[B] + I];
            val[H] = src.template load<H>();
            load_vals<B, I + 1>(val, src);
        }
    }
    template <int B, int I, class Src, int NV>
    __device__ static inline void finish_vals(zc (&val)[NV], const Src& src) {
        if constexpr (FU.start[B] + I < FU.start[B + 1]) {
            constexpr int H = FU.h[FU.start[B] + I];
            val[H] = src.finish(val[H]);
            finish_vals<B, I + 1>(val, src);
        }
    }
    template <int B, class Src, int NV>
    __device__ static inline void issue(zc (&val)[NV], zc (&eb)[D + 1], const zc* __restrict__ tab_lane,
                                        const Src& src) {
        if constexpr (B < NB) {
            constexpr int off = P::T.pbase[G] + 16 * B;
            eb[B % (D + 1)] = tab_lane[off];
            load_vals<B, 0>(val, src);
        }
    }
    template <int B, int N, class Acc, int NV>
    __device__ static inline void entries(Acc& acc, const zc (&val)[NV], const zc& e) {
        constexpr int E = S + 16 * B + N;
        if constexpr (N < 16 && E < END) {
            constexpr int slot = Slot::of(P::T.comp[E]);
            constexpr int h = P::T.hidx[E];
            zmac_bcast<N>(acc[slot], e, val[h]);
            entries<B, N + 1>(acc, val, e);
        }
    }
    template <int B, class Acc, class Src, int NV>
    __device__ static inline void batches(Acc& acc, zc (&val)[NV], zc (&eb)[D + 1], const zc* __restrict__ tab_lane,
                                          const Src& src) {
        if constexpr (B < NB) {
            issue<B + D>(val, eb, tab_lane, src);
            __builtin_amdgcn_sched_barrier(0);
            finish_vals<B, 0>(val, src);
            entries<B, 0>(acc, val, eb[B % (D + 1)]);
            __builtin_amdgcn_sched_barrier(0);
            batches<B + 1>(acc, val, eb, tab_lane, src);
        }
    }
    template <int B, class Src, int NV>
    __device__ static inline void prologue(zc (&val)[NV], zc (&eb)[D + 1], const zc* __restrict__ tab_lane,
                                           const Src& src) {
        if constexpr (B < D) {
            issue<B>(val, eb, tab_lane, src);
            prologue<B + 1>(val, eb, tab_lane, src);
        }
    }
    template <class Acc, class Src, int NV>
    __device__ static inline void run(Acc& acc, zc (&val)[NV], const zc* __restrict__ tab_lane, const Src& src) {
        zc eb[D + 1];
        prologue<0>(val, eb, tab_lane, src);
        batches<0>(acc, val, eb, tab_lane, src);
    }
};

constexpr int kKetDepth = 3;  // ket batches in flight ahead of the one computing
constexpr int kBraDepth = 1;  // bra batches (the tile values come from LDS)

template <class C, int CI>
struct KetSlot {  // ket component -> W register of chunk CI
    static constexpr int of(int c) { return c - CI * C::TILE; }
};
template <class C>
struct BraSlot {  // bra component -> OUT register of its group
    static constexpr int of(int c) { return c / kG; }
};

// R0[k + j] of the lane's row, scaled by the ket pair coefficient.
template <class C>
struct R0Src {
    const zc* __restrict__ rq;
    int roff;
    const int (&kb)[C::NTP];
    double ccd;
    template <int H>
    __device__ inline zc load() const {
        constexpr plan::P3 j = plan::herm(C::LCD, H);
        const unsigned off = (unsigned)(roff + kb[C::tp(j.x, j.y)]) * 16u + 16u * j.z;
        return *reinterpret_cast<const zc*>(reinterpret_cast<const char*>(rq) + off);
    }
    __device__ inline zc finish(zc r) const { return zc{ccd * r.x, ccd * r.y}; }
};

// W[k][cd] of the LDS tile for the lane's ket component.
template <class C>
struct WSrc {
    const zc* Wl;
    int cl;
    template <int H>
    __device__ inline zc load() const {
        return Wl[H * C::WSTRIDE + cl];
    }
    __device__ inline zc finish(zc w) const { return w; }
};

// W[cd] += sum_j E_cd[j] ccd R0[k + j] for the lane's row (bra pair, k).
template <class C, int CI>
__device__ inline void ket_step(zc (&W)[C::TILE], const zc* __restrict__ rq, int roff, const int (&kb)[C::NTP],
                                double ccd, const zc* __restrict__ ek_lane) {
    zc rv[C::NHCD];
    const R0Src<C> src{rq, roff, kb, ccd};
    Pipe<typename C::Ket, CI, kKetDepth, KetSlot<C, CI>>::run(W, rv, ek_lane, src);
}

// OUT[ab] += sum_k E_ab[k] W[k][cd] for group G's bra components (lane = cd).
template <class C, int G>
__device__ inline void bra_step(zc (&OUT)[C::NOUT], const zc* Wl, int cl, const zc* __restrict__ eb_lane) {
    zc wv[C::NHAB];
    const WSrc<C> src{Wl, cl};
    Pipe<typename C::Bra, G, kBraDepth, BraSlot<C>>::run(OUT, wv, eb_lane, src);
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
inline constexpr Norms<C> kNorms = make_norms<C>();
template <class C>
__constant__ Norms<C> c_norms = kNorms<C>;

// Stores (or, for later passes over the bra pairs, adds) the group's outputs.
template <class C, int G, int AB>
__device__ inline void write_ab(const zc (&OUT)[C::NOUT], double gcd, long long base, const FastItem& it,
                                bool accumulate, zc* __restrict__ out) {
    if constexpr (AB < C::NAB) {
        constexpr double g = kNorms<C>.ab[AB];
        constexpr int ca = AB / C::NCB, cb = AB % C::NCB;
        zc v = zscale(g * gcd, OUT[AB / kG]);
        if (it.conj) v.y = -v.y;
        const long long ia = (long long)it.x[0] * C::NCA + ca, ib = (long long)it.x[1] * C::NCB + cb;
        zc* dst = out + base + ia * it.stride[0] + ib * it.stride[1];
        *dst = accumulate ? zadd(*dst, v) : v;
        write_ab<C, G, AB + kG>(OUT, gcd, base, it, accumulate, out);
    }
}

template <class C, int G>
__device__ inline void write_out(const zc (&OUT)[C::NOUT], int lane, int c0, int nthis, const FastItem& it,
                                 bool accumulate, zc* __restrict__ out) {
    if (lane >= nthis) return;
    const int cd = c0 + opaque_v(lane);
    const int cc = cd / C::NCDD, dc = cd % C::NCDD;
    const double gcd = c_norms<C>.cd[cd];
    const long long ic = (long long)it.x[2] * C::NCC + cc, id = (long long)it.x[3] * C::NCDD + dc;
    const long long base = it.out_base + ic * it.stride[2] + id * it.stride[3];
    write_ab<C, G, G>(OUT, gcd, base, it, accumulate, out);
}

// Per-lane R0 offsets of bra Hermite index k: kb[(tau,phi)] = index(k_t + tau, k_u + phi, k_v)
// in the lexicographic table of degree <= L.
template <class C>
__device__ inline void r0_offsets(int k, int (&kb)[C::NTP]) {
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
        for (int p = 0; p + t <= C::LCD; ++p) kb[C::tp(t, p)] = plan::lexoff(C::L, kt + t, ku + p) + kv;
}

// Per-workgroup constants of one work item.
struct ItemView {
    const double *ca, *cb, *cc, *cd;  // contraction coefficients of the item's rows
    int Ka, Kb, Kc, Kd;
};

// The bra primitive pairs of a pass, one per round: the lanes holding the
// round's rows store W (scaled by the bra pair coefficient) into an LDS tile,
// then the waves with a bra group accumulate OUT from it. Two tiles alternate,
// so one barrier per round suffices. Every wave executes the same number of
// barriers.
template <class C, int CI, int G>
__device__ inline void bra_rounds(const zc (&W)[C::TILE], double cab, bool active, int prow, int k, int npab, zc* wl,
                                  int cl, int lane, const zc* __restrict__ eb, const FastItem& it, bool accumulate,
                                  zc* __restrict__ out) {
    constexpr int c0 = CI * C::TILE;
    constexpr int nthis = (C::NCD - c0) < C::TILE ? (C::NCD - c0) : C::TILE;
    zc OUT[C::NOUT];
#pragma unroll
    for (int i = 0; i < C::NOUT; ++i) OUT[i] = zc{0.0, 0.0};
    for (int r = 0; r < npab; ++r) {
        zc* buf = wl + (r & 1) * C::WBUF;
        if (active && prow == r) {
            const double c = opaque_d(cab);  // scale here, not in a hoisted copy of W
#pragma unroll
            for (int t = 0; t < nthis; ++t) buf[k * C::WSTRIDE + t] = zscale(c, W[t]);
        }
        __syncthreads();
        if constexpr (G < kG) bra_step<C, G>(OUT, buf, cl, eb + r * C::Bra::NEP);
    }
    __syncthreads();  // the tiles are rewritten by the next pass or chunk
    if constexpr (G < kG) write_out<C, G>(OUT, lane, c0, nthis, it, accumulate, out);
}

template <class C, int CI>
__device__ inline void run_chunk(const FastItem& it, const ItemView& v, const zc* __restrict__ rq,
                                 const zc* __restrict__ etab, zc* __restrict__ out, zc* wl, int wave, int lane,
                                 int prow, int k, int ppp) {
    constexpr int c0 = CI * C::TILE;
    constexpr int nthis = (C::NCD - c0) < C::TILE ? (C::NCD - c0) : C::TILE;
    const int Kab = v.Ka * v.Kb, Kcd = v.Kc * v.Kd;
    const int cl = lane < nthis ? lane : nthis - 1;
    const int l16 = lane & 15;
    const zc* ek = etab + it.eket_off + l16;
    for (int pab0 = 0; pab0 < Kab; pab0 += ppp) {
        const int npab = Kab - pab0 < ppp ? Kab - pab0 : ppp;
        const bool active = prow < npab;
        const int pab = pab0 + (active ? prow : 0);
        zc W[C::TILE];
#pragma unroll
        for (int t = 0; t < C::TILE; ++t) W[t] = zc{0.0, 0.0};
        // A wavefront with no row of this pass (the last pass of a quartet
        // whose bra pairs do not fill the workgroup) skips the ket
        // contraction: its W is never stored. The test is wave-uniform.
        if (wave * 64 < npab * C::NHAB) {
            int kb[C::NTP];  // recomputed per pass: not live through the bra rounds
            r0_offsets<C>(opaque_v(k), kb);
            int roff = pab * Kcd * C::NHL;
            for (int pcd = 0; pcd < Kcd; ++pcd, roff += C::NHL) {
                const int pc = pcd / v.Kd, pd = pcd - pc * v.Kd;
                const double ccd = v.cc[pc] * v.cd[pd];
                ket_step<C, CI>(W, rq, opaque_v(roff), kb, ccd, ek + opaque_s(pcd) * C::Ket::NEP);
            }
        }
        const int pa = pab / v.Kb, pb = pab - pa * v.Kb;
        const double cab = v.ca[pa] * v.cb[pb];
        const zc* eb = etab + it.ebra_off + (long long)pab0 * C::Bra::NEP + l16;
        // each wave runs the rounds with its own bra group; waves beyond kG only supply their rows
        switch (wave) {
            case 0: bra_rounds<C, CI, 0>(W, cab, active, prow, k, npab, wl, cl, lane, eb, it, pab0 > 0, out); break;
            case 1: bra_rounds<C, CI, 1>(W, cab, active, prow, k, npab, wl, cl, lane, eb, it, pab0 > 0, out); break;
            case 2: bra_rounds<C, CI, 2>(W, cab, active, prow, k, npab, wl, cl, lane, eb, it, pab0 > 0, out); break;
            case 3: bra_rounds<C, CI, 3>(W, cab, active, prow, k, npab, wl, cl, lane, eb, it, pab0 > 0, out); break;
            default: bra_rounds<C, CI, kG>(W, cab, active, prow, k, npab, wl, cl, lane, eb, it, pab0 > 0, out); break;
        }
    }
}

template <class C, int CI>
struct ChunkLoop {
    __device__ static inline void run(const FastItem& it, const ItemView& v, const zc* rq, const zc* etab, zc* out,
                                      zc* wl, int wave, int lane, int prow, int k, int ppp) {
        run_chunk<C, CI>(it, v, rq, etab, out, wl, wave, lane, prow, k, ppp);
        if constexpr (CI + 1 < C::NCHUNK) ChunkLoop<C, CI + 1>::run(it, v, rq, etab, out, wl, wave, lane, prow, k, ppp);
    }
};

// One workgroup (blockDim.x = 64 * waves, waves >= kG) per work item. A pass
// covers ppp = min(Kab, blockDim.x / NHAB) bra primitive pairs; row
// threadIdx.x is (bra pair threadIdx.x / NHAB of the pass, k = threadIdx.x % NHAB).
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(kMaxThreads) void fast_transform_kernel(const FastItem* __restrict__ items,
                                                                     const DevShell* __restrict__ shells,
                                                                     const double* __restrict__ coefs,
                                                                     const zc* __restrict__ r0,
                                                                     const zc* __restrict__ etab,
                                                                     zc* __restrict__ out) {
    using C = Cls<LA, LB, LC, LD>;
    __shared__ zc wl[2 * C::WBUF];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int lane = threadIdx.x % 64;
    const FastItem it = items[blockIdx.x];
    ItemView v;
    v.Ka = shells[it.sh[0]].nprim;
    v.Kb = shells[it.sh[1]].nprim;
    v.Kc = shells[it.sh[2]].nprim;
    v.Kd = shells[it.sh[3]].nprim;
    v.ca = coefs + shells[it.sh[0]].coef_off + it.x[0] * v.Ka;
    v.cb = coefs + shells[it.sh[1]].coef_off + it.x[1] * v.Kb;
    v.cc = coefs + shells[it.sh[2]].coef_off + it.x[2] * v.Kc;
    v.cd = coefs + shells[it.sh[3]].coef_off + it.x[3] * v.Kd;
    const int kab = v.Ka * v.Kb;
    const int fit = (int)blockDim.x / C::NHAB;
    const int ppp = kab < fit ? kab : fit;
    const int prow = threadIdx.x / C::NHAB, k = threadIdx.x - prow * C::NHAB;
    ChunkLoop<C, 0>::run(it, v, r0 + it.r0_off, etab, out, wl, wave, lane, prow, k, ppp);
}

// ---------------------------------------------------------------------------
// Launch table
// ---------------------------------------------------------------------------
template <int LA, int LB, int LC, int LD>
hipError_t launch_produce(const FastLaunchArgs& a, hipStream_t s) {
    using C = Cls<LA, LB, LC, LD>;
    if (a.npq) {
        pair_kernel<LA, LB, kG, 1, false><<<dim3((unsigned)a.npq), dim3(256), 0, s>>>(a.pq, 0, a.shells, a.exps,
                                                                                     a.kappa, a.etab);
        pair_kernel<LC, LD, C::NCHUNK, 0, true><<<dim3((unsigned)a.npq), dim3(256), 0, s>>>(a.pq, 1, a.shells,
                                                                                           a.exps, a.kappa, a.etab);
        if constexpr ((C::L + 1) * (C::L + 2) / 2 <= 64)
            r0_kernel<C::L><<<dim3((unsigned)a.npq), dim3(256), 0, s>>>(a.pq, a.seeds, a.r0);
        else
            r0_lds_kernel<C::L><<<dim3((unsigned)a.npq), dim3(256), 0, s>>>(a.pq, a.seeds, a.r0);
    }
    return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
hipError_t launch_transform(const FastLaunchArgs& a, hipStream_t s) {
    using C = Cls<LA, LB, LC, LD>;
    if (a.waves < kG || a.waves > kMaxWaves || 64 * a.waves < C::NHAB) return hipErrorInvalidValue;
    if (a.nitems)
        fast_transform_kernel<LA, LB, LC, LD><<<dim3((unsigned)a.nitems), dim3(64 * a.waves), 0, s>>>(
            a.items, a.shells, a.coefs, a.r0, a.etab, a.out);
    return hipGetLastError();
}

template <int LA, int LB, int LC, int LD>
FastClassInfo info() {
    using C = Cls<LA, LB, LC, LD>;
    FastClassInfo f{};
    f.available = true;
    f.bra_ne = C::Bra::NEP;
    f.ket_ne = C::Ket::NEP;
    f.nhab = C::NHAB;
    f.nhl = C::NHL;
    f.produce = &launch_produce<LA, LB, LC, LD>;
    f.transform = &launch_transform<LA, LB, LC, LD>;
    return f;
}

}  // namespace fast
}  // namespace gh
