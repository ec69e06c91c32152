// E.R.E contraction on the FP64 matrix cores (v_mfma_f64_16x16x4f64),
// templated on the total angular momenta (LAB, LCD) of the bra and ket pairs.
//
// Per primitive quartet (bra pair p, ket pair q) the R0 table at the complex
// displacement is built in LDS from the seeds, then
//   GEMM1  W_p[k][cd] += sum_j R0_pq[k + j] Ec_q[j][cd]      (k: bra Hermite,
//          j: ket Hermite, Ec_q = (-1)^|j| Et_x Et_y Et_z of ket component cd
//          times the ket pair's coefficients and normalizations)
// accumulated over the ket pairs q, and once per bra pair
//   GEMM2  V[ab][cd] += sum_k Ea_p[ab][k] W_p[k][cd].
// Both are dense complex products (the structural zeros of the Hermite
// products are multiplied, not skipped): each complex product is four real
// MFMAs, re = Ar Br + (-Ai) Bi, im = Ar Bi + Ai Br. The operands are formed
// per MFMA step in registers: A of GEMM1 gathers R0[k + j] from LDS, B of
// GEMM1 and A of GEMM2 are products of the one-dimensional Hermite tables in
// LDS, B of GEMM2 is W from LDS. MFMA lane layout (16x16x4, f64): lane l
// supplies A[l%16][l/16] and B[l/16][l%16] and holds D[l/16 + 4r][l%16] in
// register r.
//
// Split-FP32 form (SPLIT = true, the precision study of the high angular
// momentum classes): every operand is split into FP32 hi + lo parts
// (hi = fl32(x), lo = fl32(x - hi)), the products hi*hi, hi*lo and lo*hi of a
// GEMM step (K = 4) are accumulated from zero in FP32 on
// v_mfma_f32_16x16x4f32 (hi*hi and the cross terms in separate
// accumulators), and each step's sums are added to FP64 accumulators
// (compensated into FP64). FP32's range is kept by powers of two: the R0
// seeds of a primitive quartet are scaled by 2^-e (e the largest exponent of
// s_0..s_L) and W of a bra pair by 2^-e' (e' the exponent of max |W|), undone
// exactly when a step's sums enter the FP64 accumulators; so are the bra and
// ket factors (contraction coefficients times normalizations) of a pair. Output lane
// layout of the FP32 instruction: lane l holds D[4(l/16) + r][l%16].
//
// One workgroup of 8 wavefronts takes one work item: a shell quartet, one
// combination of contraction rows and a range of at most 32 ket components
// (two 16-wide n-tiles; 16 components, one n-tile, for bra momentum 7, 8).
// GEMM tiles (m-tile, n-tile) go round robin to the wavefronts, m-tile
// major, so with two n-tiles every wavefront keeps one n-tile (its B operand
// is formed once per step for all its tiles).
#pragma once

#include "gh_hermite.h"
#include "gh_internal.h"
#include "gh_plan.h"

#include <type_traits>

namespace gh {
namespace mf {

typedef double mf4 __attribute__((ext_vector_type(4)));
typedef float mf4f __attribute__((ext_vector_type(4)));

constexpr int kWaves = 8;
constexpr int kNT = 64 * kWaves;

// f(std::integral_constant<int, n>) for a wave-uniform n in 0..N: the GEMM
// step loops run over a compile-time number of owned tiles, so no branch
// around a matrix instruction sits inside them.
template <int N, class F>
__device__ inline void with_tile_count(int n, F&& f) {
    if constexpr (N == 0) {
        f(std::integral_constant<int, 0>{});
    } else {
        if (n == N) f(std::integral_constant<int, N>{});
        else with_tile_count<N - 1>(n, f);
    }
}

// Ket n-tiles (16 components each) of one work item: two, or one where the
// bra side is so large that the accumulators of two would not fit in the
// registers of 8 wavefronts.
constexpr int n_tiles(int LAB) { return LAB >= 7 ? 1 : 2; }

constexpr int maxi(int a, int b) { return a > b ? a : b; }

// Largest component count and (l1 + 1)(l2 + 1) over shell pairs of total
// momentum LP with both shells at most kMaxShellL.
constexpr int max_pair_comps(int LP) {
    int m = 1;
    for (int l1 = 0; l1 <= LP; ++l1)
        if (l1 <= kMaxShellL && LP - l1 <= kMaxShellL) m = maxi(m, plan::nc(l1) * plan::nc(LP - l1));
    return m;
}
constexpr int max_pair_rows(int LP) {
    int m = 1;
    for (int l1 = 0; l1 <= LP; ++l1)
        if (l1 <= kMaxShellL && LP - l1 <= kMaxShellL) m = maxi(m, (l1 + 1) * (LP - l1 + 1));
    return m;
}

// Compile-time variants of an instance (VAR bits), chosen at run time by the
// autotuner: bit 0 flips the GEMM step unroll (1 <-> 2), bit 1 pads the row
// stride of the W and Ec tiles in LDS by one complex entry (16 bytes), which
// moves consecutive rows to other banks. Neither changes any sum.
static_assert(kMfmaVariantCount == 4, "variants 0..3");

template <int LAB, int LCD, int VAR = 0>
struct Shape {
    static constexpr int L = LAB + LCD;
    static constexpr int NHAB = plan::nh(LAB), NHCD = plan::nh(LCD), NHL = plan::nh(L);
    static constexpr int NTI = n_tiles(LAB), CDT = 16 * NTI;      // n-tiles, ket components per item
    static constexpr int MT1 = (NHAB + 15) / 16;                  // GEMM1 m-tiles (bra Hermite rows)
    static constexpr int KS1 = (NHCD + 3) / 4;                    // GEMM1 steps over the ket Hermite index
    static constexpr int MT2 = (max_pair_comps(LAB) + 15) / 16;   // GEMM2 m-tiles (bra components)
    static constexpr int KS2 = MT1 * 4;                           // GEMM2 steps over the bra Hermite index
    static constexpr int TW1 = (NTI * MT1 + kWaves - 1) / kWaves;  // GEMM1 tiles per wavefront
    static constexpr int TW2 = (NTI * MT2 + kWaves - 1) / kWaves;  // GEMM2 tiles per wavefront
    static constexpr int RB = max_pair_rows(LAB), RK = max_pair_rows(LCD);
    static constexpr int E1B = 3 * RB * (LAB + 1), E1K = 3 * RK * (LCD + 1);
    static constexpr int WROWS = 16 * MT1, JROWS = 4 * (KS1 + 1);  // one zero step past the end (prefetch)
    static constexpr int RPT = (NHL + kNT - 1) / kNT;             // R0 entries per thread
    static constexpr int UNROLL0 = TW1 * TW2 > 2 ? 1 : 2;         // GEMM step unroll (register budget)
    static constexpr int UNROLL = (VAR & 1) ? 3 - UNROLL0 : UNROLL0;
    static constexpr int CDS = CDT + ((VAR & 2) ? 1 : 0);          // row stride of W and Ec in LDS
    // LDS, complex entries: two R0 level buffers, the bra and ket 1-D
    // tables, W, the dense ket products Ec[j][cd]; then ints: lexicographic
    // offsets, the bra Hermite decode
    static constexpr int Z_R0 = 0, Z_E1B = 2 * (NHL + 1), Z_E1K = Z_E1B + E1B, Z_W = Z_E1K + E1K,
                         Z_EC = Z_W + (WROWS + 4) * CDS, Z_FCOL = Z_EC + JROWS * CDS, Z_SEED = Z_FCOL + CDT,
                         Z_END = Z_SEED + L + 4;
    // ints: lexicographic offsets (t * (L + 1) + u -> lex of (t, u, 0)); the
    // ket Hermite indices as (t * (L + 1) + u) | v << 16; the bra Hermite
    // indices as t | u << 8 | v << 16; per ket column its three 1-D table rows
    static constexpr int I_OFFL = 0, I_HCD = (L + 1) * (L + 1), I_HAB = I_HCD + JROWS,
                         I_COL = I_HAB + 4 * KS2 + 4, I_TUV = I_COL + 3 * CDT, I_EXP = I_TUV + NHCD,
                         I_END = I_EXP + kWaves + 1;  // split form: per-wave W exponents, the seed exponent
    static constexpr int LDS_BYTES = Z_END * 16 + I_END * 4;
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
__constant__ RList<L> c_rl = make_rlist<L>();

// Hermite indices of degree <= D in lexicographic order: t | u << 8 | v << 16.
template <int D>
struct HTab {
    int e[plan::nh(D)];
};
template <int D>
constexpr HTab<D> make_htab() {
    HTab<D> h{};
    int k = 0;
    for (int t = 0; t <= D; ++t)
        for (int u = 0; u + t <= D; ++u)
            for (int v = 0; v + u + t <= D; ++v) h.e[k++] = t | (u << 8) | (v << 16);
    return h;
}
template <int D>
__constant__ HTab<D> c_htab = make_htab<D>();

// x-major Cartesian components of shells through kMaxShellL: x | y << 8 | z << 16.
struct CartTable {
    unsigned e[plan::nh(kMaxShellL)];
};
constexpr CartTable make_carts() {
    CartTable t{};
    int k = 0;
    for (int l = 0; l <= kMaxShellL; ++l)
        for (int c = 0; c < plan::nc(l); ++c) {
            const plan::P3 p = plan::cart(l, c);
            t.e[k++] = (unsigned)p.x | ((unsigned)p.y << 8) | ((unsigned)p.z << 16);
        }
    return t;
}
__constant__ inline CartTable c_carts = make_carts();
__device__ inline unsigned cart3(int l, int c) { return c_carts.e[l * (l + 1) * (l + 2) / 6 + c]; }

// Complex 1-D Hermite table of one pair on one axis into row[0 .. D]: the
// coefficients up to i + j (built in registers), zeros above (the dense
// products read every t).
template <int D>
__device__ inline void hermite_row(int i, int j, double PA, double PB, double p, double q, zc* row) {
    zc tmp[D + 1];
    phase_hermite_regs<D + 1>(i, j, PA, PB, p, q, tmp);
#pragma unroll
    for (int t = 0; t <= D; ++t) row[t] = t <= i + j ? tmp[t] : zc{0.0, 0.0};
}

// Row of the accumulator register r of lane group lk (lane / 16): the FP64
// instruction holds D[lk + 4r][.], the FP32 one D[4 lk + r][.].
template <bool SPLIT>
__device__ constexpr int drow(int lk, int r) {
    return SPLIT ? 4 * lk + r : lk + 4 * r;
}

__device__ inline float hi32(double x) { return (float)x; }
__device__ inline float lo32(double x, float h) { return (float)(x - (double)h); }

// One GEMM step (K = 4) of a complex product in split FP32, added to the FP64
// accumulators times 2^scale_exp (ldexp: exact, and safe where the power of
// two alone would overflow or underflow).
__device__ inline void split_step(mf4& re, mf4& im, const zc a, const zc b, const int scale_exp) {
    const float arh = hi32(a.x), arl = lo32(a.x, arh), aih = hi32(a.y), ail = lo32(a.y, aih);
    const float brh = hi32(b.x), brl = lo32(b.x, brh), bih = hi32(b.y), bil = lo32(b.y, bih);
    const mf4f z = {0.0f, 0.0f, 0.0f, 0.0f};
    mf4f rh = __builtin_amdgcn_mfma_f32_16x16x4f32(arh, brh, z, 0, 0, 0);
    mf4f ih = __builtin_amdgcn_mfma_f32_16x16x4f32(arh, bih, z, 0, 0, 0);
    mf4f rx = __builtin_amdgcn_mfma_f32_16x16x4f32(arh, brl, z, 0, 0, 0);
    mf4f ix = __builtin_amdgcn_mfma_f32_16x16x4f32(arh, bil, z, 0, 0, 0);
    rh = __builtin_amdgcn_mfma_f32_16x16x4f32(-aih, bih, rh, 0, 0, 0);
    ih = __builtin_amdgcn_mfma_f32_16x16x4f32(aih, brh, ih, 0, 0, 0);
    rx = __builtin_amdgcn_mfma_f32_16x16x4f32(arl, brh, rx, 0, 0, 0);
    ix = __builtin_amdgcn_mfma_f32_16x16x4f32(arl, bih, ix, 0, 0, 0);
    rx = __builtin_amdgcn_mfma_f32_16x16x4f32(-aih, bil, rx, 0, 0, 0);
    ix = __builtin_amdgcn_mfma_f32_16x16x4f32(aih, brl, ix, 0, 0, 0);
    rx = __builtin_amdgcn_mfma_f32_16x16x4f32(-ail, bih, rx, 0, 0, 0);
    ix = __builtin_amdgcn_mfma_f32_16x16x4f32(ail, brh, ix, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        re[r] += ldexp((double)rh[r] + (double)rx[r], scale_exp);
        im[r] += ldexp((double)ih[r] + (double)ix[r], scale_exp);
    }
}

// Exponent of the largest magnitude over a wavefront (lanes pass m >= 0),
// or INT_MIN / 2 when every lane passes zero.
__device__ inline int wave_max_exponent(double m) {
    int e = m > 0.0 ? ilogb(m) : -(1 << 30);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) e = max(e, __shfl_xor(e, o));
    return e;
}

template <int LAB, int LCD, bool SPLIT = false, int VAR = 0>
__global__ __launch_bounds__(kNT) void mfma_transform_kernel(const WorkItem* __restrict__ items,
                                                             const DevShell* __restrict__ shells,
                                                             const double* __restrict__ exps,
                                                             const double* __restrict__ coefs,
                                                             const double* __restrict__ norms,
                                                             const double* __restrict__ kappa,
                                                             const zc* __restrict__ seeds, zc* __restrict__ out) {
    using S = Shape<LAB, LCD, VAR>;
    constexpr int L = S::L;
    extern __shared__ __align__(16) unsigned char smem[];
    zc* zs = reinterpret_cast<zc*>(smem);
    int* is = reinterpret_cast<int*>(smem + S::Z_END * 16);
    zc* r0buf[2] = {zs + S::Z_R0, zs + S::Z_R0 + S::NHL + 1};  // each with a zero entry at NHL
    zc* e1b = zs + S::Z_E1B;
    zc* e1k = zs + S::Z_E1K;
    zc* Wl = zs + S::Z_W;
    zc* Ec = zs + S::Z_EC;
    double* fcol = reinterpret_cast<double*>(zs + S::Z_FCOL);  // ket column factors of the current ket pair
    zc* sseed = zs + S::Z_SEED;                                 // X, s_0 .. s_L of the current primitive quartet
    int* offL = is + S::I_OFFL;
    int* hcd = is + S::I_HCD;
    int* hab = is + S::I_HAB;
    int* colrow = is + S::I_COL;
    int* tuvcd = is + S::I_TUV;  // ket Hermite indices as t | u << 8 | v << 16
    int* wexp = is + S::I_EXP;   // split form: W exponent per wavefront, then the seed exponent

    const WorkItem it = items[blockIdx.x];
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64), lane = tid % 64;
    const int li = lane % 16, lk = lane / 16;
    const DevShell& A = shells[it.sh[0]];
    const DevShell& B = shells[it.sh[1]];
    const DevShell& C = shells[it.sh[2]];
    const DevShell& D = shells[it.sh[3]];
    const int la = A.l, lb = B.l, lc = C.l, ld = D.l;
    const int ncb = ncart(lb), ncd_d = ncart(ld);
    const int nab = ncart(la) * ncb;
    const int ncomp = it.cd1 - it.cd0;  // ket components of the item (<= S::CDT)
    const int ntv = (S::NTI == 2 && ncomp > 16) ? 2 : 1;
    const int Kb = B.nprim, Kd = D.nprim;
    const int Kab = A.nprim * Kb, Kcd = C.nprim * Kd;
    const int rb_i = lb + 1, rk_i = ld + 1;  // 1-D table rows: (i * (l2 + 1) + j), per axis

    for (int e = tid; e < (L + 1) * (L + 1); e += kNT) {
        const int t = e / (L + 1), u = e % (L + 1);
        offL[e] = t + u <= L ? plan::lexoff(L, t, u) : 0;
    }
    for (int e = tid; e < 4 * S::KS2 + 4; e += kNT) hab[e] = e < S::NHAB ? c_htab<LAB>.e[e] : 0;
    if (tid < 2) r0buf[tid][S::NHL] = zc{0.0, 0.0};
    for (int e = tid; e < S::JROWS; e += kNT) {
        const int h = e < S::NHCD ? c_htab<LCD>.e[e] : 0;
        hcd[e] = ((h & 255) * (L + 1) + ((h >> 8) & 255)) | ((h >> 16) << 16);
        if (e < S::NHCD) tuvcd[e] = h;
    }
    for (int e = tid; e < S::CDT; e += kNT) {
        int rx = 0, ry = 0, rz = 0;
        if (e < ncomp) {
            const int cdc = it.cd0 + e;
            const unsigned cp = cart3(lc, cdc / ncd_d), dp = cart3(ld, cdc % ncd_d);
            rx = ((cp & 255) * rk_i + (dp & 255)) * (LCD + 1);
            ry = (S::RK + ((cp >> 8) & 255) * rk_i + ((dp >> 8) & 255)) * (LCD + 1);
            rz = (2 * S::RK + (cp >> 16) * rk_i + (dp >> 16)) * (LCD + 1);
        }
        colrow[e] = rx;
        colrow[S::CDT + e] = ry;
        colrow[2 * S::CDT + e] = rz;
    }

    // The R0 recurrence steps of this thread's entries, fixed for the item:
    // rs_a = lex | src1 << 16, rs_b = src2 | axis << 16 | coefficient << 18
    // (axis 3: the seed entry; sources index the previous level lexicographically).
    int rs_a[S::RPT], rs_b[S::RPT];
#pragma unroll
    for (int r = 0; r < S::RPT; ++r) {
        const int e = tid + r * kNT;
        rs_a[r] = 0;
        rs_b[r] = 3 << 16;
        if (e < S::NHL) {
            const int pk = c_rl<L>.e[e];
            const int t = pk & 63, u = (pk >> 6) & 63, v = (pk >> 12) & 63, lex = pk >> 18;
            int s1 = 0, s2 = 0, ax = 3, cf = 0;
            if (t > 0) {
                ax = 0;
                s1 = plan::lexoff(L, t - 1, u) + v;
                if (t > 1) s2 = plan::lexoff(L, t - 2, u) + v, cf = t - 1;
            } else if (u > 0) {
                ax = 1;
                s1 = plan::lexoff(L, 0, u - 1) + v;
                if (u > 1) s2 = plan::lexoff(L, 0, u - 2) + v, cf = u - 1;
            } else if (v > 0) {
                ax = 2;
                s1 = v - 1;
                if (v > 1) s2 = v - 2, cf = v - 1;
            }
            rs_a[r] = lex | (s1 << 16);
            rs_b[r] = s2 | (ax << 16) | (cf << 18);
        }
    }

    // GEMM1 tiles of this wavefront: m-tile major, t = wave + kWaves * i
    int g1_mt[S::TW1], g1_k[S::TW1];
    bool g1_on[S::TW1], g1_row[S::TW1];  // tile owned (wave-uniform); the lane's row k < NHAB
#pragma unroll
    for (int i = 0; i < S::TW1; ++i) {
        const int t = wave + kWaves * i;
        g1_on[i] = __builtin_amdgcn_readfirstlane((int)(t < S::MT1 * ntv)) != 0;
        g1_mt[i] = t / ntv;
        const int k = g1_mt[i] * 16 + li;
        g1_row[i] = k < S::NHAB;
        const int h = k < S::NHAB ? c_htab<LAB>.e[k] : 0;
        g1_k[i] = ((h & 255) * (L + 1) + ((h >> 8) & 255)) | ((h >> 16) << 16);
    }
    int n1 = 0;  // owned GEMM1 tiles: a prefix of the TW1 slots (t grows with i)
#pragma unroll
    for (int i = 0; i < S::TW1; ++i) n1 += g1_on[i] ? 1 : 0;
    const int nt1 = ntv == 2 ? (wave & 1) : 0;  // the wavefront's n-tile (fixed: kWaves is even)
    const int cdl = nt1 * 16 + li;              // the lane's ket column (GEMM B operands, output)
    const bool cd_on = cdl < ncomp;
    int cc = 0, dc = 0;
    if (cd_on) {
        const int cdc = it.cd0 + cdl;
        cc = cdc / ncd_d;
        dc = cdc % ncd_d;
    }

    // GEMM2 tiles of this wavefront (the same n-tile)
    int g2_mt[S::TW2];
    bool g2_on[S::TW2];
    int g2_bx[S::TW2], g2_by[S::TW2], g2_bz[S::TW2];  // 1-D table rows of the lane's bra component
    int g2_ab[S::TW2];
#pragma unroll
    for (int i = 0; i < S::TW2; ++i) {
        const int t = wave + kWaves * i;
        g2_on[i] = __builtin_amdgcn_readfirstlane((int)(t < S::MT2 * ntv && (t / ntv) * 16 < nab)) != 0;
        g2_mt[i] = t / ntv;
        const int ab = g2_mt[i] * 16 + li;
        g2_ab[i] = ab < nab ? ab : -1;
        unsigned apk = 0, bpk = 0;
        if (ab < nab) {
            apk = cart3(la, ab / ncb);
            bpk = cart3(lb, ab % ncb);
        }
        g2_bx[i] = ((apk & 255) * rb_i + (bpk & 255)) * (LAB + 1);
        g2_by[i] = (S::RB + ((apk >> 8) & 255) * rb_i + ((bpk >> 8) & 255)) * (LAB + 1);
        g2_bz[i] = (2 * S::RB + (apk >> 16) * rb_i + (bpk >> 16)) * (LAB + 1);
    }

    int n2 = 0;  // owned GEMM2 tiles, likewise a prefix
#pragma unroll
    for (int i = 0; i < S::TW2; ++i) n2 += g2_on[i] ? 1 : 0;

    mf4 vre[S::TW2], vim[S::TW2];
#pragma unroll
    for (int i = 0; i < S::TW2; ++i) vre[i] = vim[i] = mf4{0.0, 0.0, 0.0, 0.0};

    const double *Ac = A.center, *Bc = B.center, *Cc = C.center, *Dc = D.center;
    for (int pab = 0; pab < Kab; ++pab) {
        const int pa = pab / Kb, pb = pab % Kb;
        const double ea = exps[A.prim_off + pa], eb = exps[B.prim_off + pb];
        const double p = ea + eb;
        __syncthreads();  // the previous bra pair's tables and W are no longer read
        for (int e = tid; e < 3 * (la + 1) * (lb + 1); e += kNT) {
            const int axis = e / ((la + 1) * (lb + 1));
            const int i = (e / (lb + 1)) % (la + 1), j = e % (lb + 1);
            const double Pc = (ea * Ac[axis] + eb * Bc[axis]) / p;
            const double qv = kappa[3 * it.sh[1] + axis] - kappa[3 * it.sh[0] + axis];
            hermite_row<LAB>(i, j, Pc - Ac[axis], Pc - Bc[axis], p, qv,
                             e1b + (axis * S::RB + i * rb_i + j) * (LAB + 1));
        }
        // the lane's bra factors (coefficients x normalizations) for this pair;
        // the split form scales them by 2^-e_ab, e_ab the exponent of the
        // pair's factor for its first components (the same in every lane)
        int e_ab = 0;
        if constexpr (SPLIT) {
            const double f0 = fabs(coefs[A.coef_off + it.x[0] * A.nprim + pa] * norms[A.norm_off + pa] *
                                   coefs[B.coef_off + it.x[1] * B.nprim + pb] * norms[B.norm_off + pb]);
            e_ab = f0 > 0.0 ? ilogb(f0) : 0;
        }
        double fab[S::TW2];
#pragma unroll
        for (int i = 0; i < S::TW2; ++i) {
            fab[i] = 0.0;
            if (g2_ab[i] >= 0) {
                const int ca = g2_ab[i] / ncb, cb = g2_ab[i] % ncb;
                fab[i] = coefs[A.coef_off + it.x[0] * A.nprim + pa] * norms[A.norm_off + ca * A.nprim + pa] *
                         coefs[B.coef_off + it.x[1] * B.nprim + pb] * norms[B.norm_off + cb * B.nprim + pb];
                if constexpr (SPLIT) fab[i] = ldexp(fab[i], -e_ab);
            }
        }

        mf4 wre[S::TW1], wim[S::TW1];
#pragma unroll
        for (int i = 0; i < S::TW1; ++i) wre[i] = wim[i] = mf4{0.0, 0.0, 0.0, 0.0};

        // Seeds (threads below L + 4) and ket column factors (threads below
        // CDT) of a primitive quartet, loaded into registers one ket pair ahead
        // so that their latency overlaps the previous pair's matrix products.
        auto load_seed = [&](int pq) -> zc {
            return tid < L + 4 ? seeds[it.seed_off + (long long)pq * (L + 4) + tid] : zc{0.0, 0.0};
        };
        auto load_factor = [&](int pcd) -> double {
            if (tid >= ncomp) return 0.0;
            const int pc = pcd / Kd, pd = pcd % Kd;
            const int cdc = it.cd0 + tid, c_ = cdc / ncd_d, d_ = cdc % ncd_d;
            return coefs[C.coef_off + it.x[2] * C.nprim + pc] * norms[C.norm_off + c_ * C.nprim + pc] *
                   coefs[D.coef_off + it.x[3] * D.nprim + pd] * norms[D.norm_off + d_ * D.nprim + pd];
        };
        zc seed_next = load_seed(pab * Kcd);
        double f_next = load_factor(0);
        for (int pcd = 0; pcd < Kcd; ++pcd) {
            const int pc = pcd / Kd, pd = pcd % Kd;
            const double ec = exps[C.prim_off + pc], ed = exps[D.prim_off + pd];
            const double q = ec + ed;
            __syncthreads();  // the previous ket pair's tables, R0 and products are no longer read
            if constexpr (SPLIT) {
                // seeds s_n scaled by 2^-e, e their largest exponent, and the ket
                // column factors by 2^-e_c, e_c theirs (wavefront 0 holds them all)
                if (wave == 0) {
                    const bool sn = tid >= 3 && tid < L + 4;
                    int e = wave_max_exponent(sn ? fmax(fabs(seed_next.x), fabs(seed_next.y)) : 0.0);
                    if (e < -(1 << 29)) e = 0;
                    int e_c = wave_max_exponent(fabs(f_next));
                    if (e_c < -(1 << 29)) e_c = 0;
                    if (tid < L + 4)
                        sseed[tid] = sn ? zc{ldexp(seed_next.x, -e), ldexp(seed_next.y, -e)} : seed_next;
                    if (tid < S::CDT) fcol[tid] = ldexp(f_next, -e_c);
                    if (tid == 0) wexp[kWaves] = e + e_c;
                }
            } else {
                if (tid < L + 4) sseed[tid] = seed_next;
                if (tid < S::CDT) fcol[tid] = f_next;
            }
            for (int e = tid; e < 3 * (lc + 1) * (ld + 1); e += kNT) {
                const int axis = e / ((lc + 1) * (ld + 1));
                const int i = (e / (ld + 1)) % (lc + 1), j = e % (ld + 1);
                const double Qc = (ec * Cc[axis] + ed * Dc[axis]) / q;
                const double qv = kappa[3 * it.sh[3] + axis] - kappa[3 * it.sh[2] + axis];
                hermite_row<LCD>(i, j, Qc - Cc[axis], Qc - Dc[axis], q, qv,
                                 e1k + (axis * S::RK + i * rk_i + j) * (LCD + 1));
            }
            __syncthreads();  // seeds and factors in LDS
            const zc X0 = sseed[0], X1 = sseed[1], X2 = sseed[2];
            // R0 recurrence: level L is s_L, level n from level n + 1
            if (tid == 0) r0buf[L & 1][0] = sseed[3 + L];
#pragma unroll 1
            for (int n = L - 1; n >= 0; --n) {
                __syncthreads();
                const zc* cur = r0buf[(n + 1) & 1];
                zc* nxt = r0buf[n & 1];
                const int cnt = plan::nh(L - n);
#pragma unroll
                for (int r = 0; r < S::RPT; ++r) {
                    if (tid + r * kNT < cnt) {
                        const int lex = rs_a[r] & 0xffff, s1 = rs_a[r] >> 16;
                        const int s2 = rs_b[r] & 0xffff, ax = (rs_b[r] >> 16) & 3, cf = rs_b[r] >> 18;
                        zc val;
                        if (ax == 3) {
                            val = sseed[3 + n];
                        } else {
                            val = zmul(ax == 0 ? X0 : (ax == 1 ? X1 : X2), cur[s1]);
                            if (cf > 0) {
                                const zc r2 = cur[s2];
                                val.x = fma((double)cf, r2.x, val.x);
                                val.y = fma((double)cf, r2.y, val.y);
                            }
                        }
                        nxt[lex] = val;
                    }
                }
            }
            __syncthreads();  // R0 and the 1-D ket tables are complete
            // dense ket products Ec[j][cd] = (-1)^|j| Et_x Et_y Et_z x factor, zero
            // past the ket Hermite range (padding and one prefetch step)
            for (int e = tid; e < S::JROWS * S::CDT; e += kNT) {
                const int j = e / S::CDT, col = e % S::CDT;
                zc v{0.0, 0.0};
                if (j < S::NHCD && col < ncomp) {
                    const int h = tuvcd[j];
                    const int tj = h & 255, uj = (h >> 8) & 255, vj = h >> 16;
                    const zc ex = e1k[colrow[col] + tj];
                    const zc ey = e1k[colrow[S::CDT + col] + uj];
                    const zc ez = e1k[colrow[2 * S::CDT + col] + vj];
                    const double f = ((tj + uj + vj) & 1) ? -fcol[col] : fcol[col];
                    v = zscale(f, zmul(zmul(ex, ey), ez));
                }
                Ec[j * S::CDS + col] = v;
            }
            __syncthreads();
            const zc* R0 = r0buf[0];
            const int sc1 = SPLIT ? wexp[kWaves] : 0;  // exponent that undoes the seed and ket scaling
            // GEMM1, one step of operands loaded ahead; the two products of
            // every tile that share no accumulator are issued back to back
            if (pcd + 1 < Kcd) {
                seed_next = load_seed(pab * Kcd + pcd + 1);
                f_next = load_factor(pcd + 1);
            }
            // the gather index of R0[k + j]; rows past NHAB read the zero entry
            auto gather = [&](int i, int hj) -> zc {
                const int idx = offL[(g1_k[i] & 0xffff) + (hj & 0xffff)] + (g1_k[i] >> 16) + (hj >> 16);
                return R0[g1_row[i] ? idx : S::NHL];
            };
            with_tile_count<S::TW1>(n1, [&](auto na_c) {
                constexpr int NA = decltype(na_c)::value;
                if constexpr (NA > 0) {
                    zc b = Ec[lk * S::CDS + cdl];
                    zc a[NA];
                    {
                        const int hj = hcd[lk];
#pragma unroll
                        for (int i = 0; i < NA; ++i) a[i] = gather(i, hj);
                    }
#pragma unroll S::UNROLL
                    for (int s = 0; s < S::KS1; ++s) {
                        const int jn = 4 * (s + 1) + lk;  // step s + 1 (a zero row past the end)
                        const zc bn = Ec[jn * S::CDS + cdl];
                        const int hjn = hcd[jn];
                        zc an[NA];
#pragma unroll
                        for (int i = 0; i < NA; ++i) an[i] = gather(i, hjn);
                        if constexpr (SPLIT) {
#pragma unroll
                            for (int i = 0; i < NA; ++i) split_step(wre[i], wim[i], a[i], b, sc1);
                        } else {
#pragma unroll
                            for (int i = 0; i < NA; ++i) {
                                wre[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b.x, wre[i], 0, 0, 0);
                                wim[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].x, b.y, wim[i], 0, 0, 0);
                            }
#pragma unroll
                            for (int i = 0; i < NA; ++i) {
                                wre[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[i].y, b.y, wre[i], 0, 0, 0);
                                wim[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i].y, b.x, wim[i], 0, 0, 0);
                            }
                        }
                        b = bn;
#pragma unroll
                        for (int i = 0; i < NA; ++i) a[i] = an[i];
                    }
                }
            });
        }
        __syncthreads();  // every wavefront finished GEMM1 reading R0
        // W to LDS: tile i holds rows mt*16 + drow(lk, r), column nt1*16 + li
#pragma unroll
        for (int i = 0; i < S::TW1; ++i) {
            if (g1_on[i]) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    Wl[(g1_mt[i] * 16 + drow<SPLIT>(lk, r)) * S::CDS + cdl] = zc{wre[i][r], wim[i][r]};
            }
        }
        if constexpr (SPLIT) {
            double m = 0.0;
#pragma unroll
            for (int i = 0; i < S::TW1; ++i)
                if (g1_on[i])
#pragma unroll
                    for (int r = 0; r < 4; ++r) m = fmax(m, fmax(fabs(wre[i][r]), fabs(wim[i][r])));
            const int e = wave_max_exponent(m);
            if (lane == 0) wexp[wave] = e;
        }
        __syncthreads();
        // split form: W enters the FP32 products scaled by 2^-e', e' the exponent of max |W|
        int w_in = 0, w_out = 0;
        if constexpr (SPLIT) {
            int e = wexp[0];
#pragma unroll
            for (int w8 = 1; w8 < kWaves; ++w8) e = max(e, wexp[w8]);
            if (e < -(1 << 29)) e = 0;
            w_in = -e;
            w_out = e + e_ab;  // undoes the W and bra factor scaling
        }
        // GEMM2, one step of operands loaded ahead (W has a zero step past the end)
        // rows k >= NHAB of W are zero, invalid bra components have fab = 0
        auto bra_operand = [&](int k, int i) -> zc {
            const int hk = hab[k];
            const zc e = zmul(zmul(e1b[g2_bx[i] + (hk & 255)], e1b[g2_by[i] + ((hk >> 8) & 255)]),
                              e1b[g2_bz[i] + (hk >> 16)]);
            return zscale(fab[i], e);
        };
        with_tile_count<S::TW2>(n2, [&](auto na_c) {
            constexpr int NA = decltype(na_c)::value;
            if constexpr (NA > 0) {
                zc w = Wl[lk * S::CDS + cdl];
                if constexpr (SPLIT) w = zc{ldexp(w.x, w_in), ldexp(w.y, w_in)};
                zc a2[NA];
#pragma unroll
                for (int i = 0; i < NA; ++i) a2[i] = bra_operand(lk, i);
#pragma unroll S::UNROLL
                for (int s = 0; s < S::KS2; ++s) {
                    const int kn = 4 * (s + 1) + lk;
                    zc wn = Wl[kn * S::CDS + cdl];
                    if constexpr (SPLIT) wn = zc{ldexp(wn.x, w_in), ldexp(wn.y, w_in)};
                    zc an[NA];
#pragma unroll
                    for (int i = 0; i < NA; ++i) an[i] = bra_operand(kn, i);
                    if constexpr (SPLIT) {
#pragma unroll
                        for (int i = 0; i < NA; ++i) split_step(vre[i], vim[i], a2[i], w, w_out);
                    } else {
#pragma unroll
                        for (int i = 0; i < NA; ++i) {
                            vre[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[i].x, w.x, vre[i], 0, 0, 0);
                            vim[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[i].x, w.y, vim[i], 0, 0, 0);
                        }
#pragma unroll
                        for (int i = 0; i < NA; ++i) {
                            vre[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a2[i].y, w.y, vre[i], 0, 0, 0);
                            vim[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2[i].y, w.x, vim[i], 0, 0, 0);
                        }
                    }
                    w = wn;
#pragma unroll
                    for (int i = 0; i < NA; ++i) a2[i] = an[i];
                }
            }
        });
    }
    // write: tile i holds bra components mt*16 + drow(lk, r), ket component cdl
    if (cd_on) {
        const long long ic = (long long)it.x[2] * ncart(lc) + cc, id = (long long)it.x[3] * ncd_d + dc;
        const long long base = it.out_base + ic * it.stride[2] + id * it.stride[3];
#pragma unroll
        for (int i = 0; i < S::TW2; ++i) {
            if (g2_on[i]) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int ab = g2_mt[i] * 16 + drow<SPLIT>(lk, r);
                    if (ab < nab) {
                        const long long ia = (long long)it.x[0] * ncart(la) + ab / ncb,
                                        ib = (long long)it.x[1] * ncb + ab % ncb;
                        out[base + ia * it.stride[0] + ib * it.stride[1]] = zc{vre[i][r], vim[i][r]};
                    }
                }
            }
        }
    }
}

}  // namespace mf
}  // namespace gh

namespace gh {
namespace mf {

// Launch of one (LAB, LCD) instance over `n` items (dynamic LDS of
// Shape::LDS_BYTES; above 64 KB the kernel's limit is raised once).
template <int LAB, int LCD, bool SPLIT = false, int VAR = 0>
hipError_t launch_one(const WorkItem* items, int n, const DevShell* shells, const double* exps,
                      const double* coefs, const double* norms, const double* kappa, const zc* seeds, zc* out,
                      hipStream_t s) {
    using S = Shape<LAB, LCD, VAR>;
    static_assert(S::LDS_BYTES <= 160 * 1024, "MFMA transform exceeds the LDS");
    if (n == 0) return hipSuccess;
    if constexpr (S::LDS_BYTES > 64 * 1024) {
        static bool raised = false;
        if (!raised) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&mfma_transform_kernel<LAB, LCD, SPLIT, VAR>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, S::LDS_BYTES);
            if (e != hipSuccess) return e;
            raised = true;
        }
    }
    mfma_transform_kernel<LAB, LCD, SPLIT, VAR><<<dim3((unsigned)n), dim3(kNT), S::LDS_BYTES, s>>>(
        items, shells, exps, coefs, norms, kappa, seeds, out);
    return hipGetLastError();
}

// Launch of one (LAB, LCD) instance in compile-time variant `var`
// (0..kMfmaVariants-1; the split form has variant 0 only).
template <int LAB, int LCD, bool SPLIT = false>
hipError_t launch_instance(int var, const WorkItem* items, int n, const DevShell* shells, const double* exps,
                           const double* coefs, const double* norms, const double* kappa, const zc* seeds, zc* out,
                           hipStream_t s) {
    if constexpr (SPLIT) {
        return launch_one<LAB, LCD, true, 0>(items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
    } else {
        switch (var) {
            case 1: return launch_one<LAB, LCD, false, 1>(items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
            case 2: return launch_one<LAB, LCD, false, 2>(items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
            case 3: return launch_one<LAB, LCD, false, 3>(items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
            default: return launch_one<LAB, LCD, false, 0>(items, n, shells, exps, coefs, norms, kappa, seeds, out, s);
        }
    }
}

}  // namespace mf
}  // namespace gh
