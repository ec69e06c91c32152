// gfx950 kernels of the London-orbital McMurchie-Davidson ERI path.
//
// Formulation. The reference (eri.cpp:283-368) builds, for every Cartesian
// component quartet, a six-index complex table
//   R^n_{h1,h2} = d_P^{h1} d_Q^{h2} [exp(-i q1.P - i q2.Q) G^n(P'-Q')]
// and contracts it with real Hermite coefficients E_ab[h1] E_cd[h2]. Expanding
// the derivatives of the phase binomially gives
//   sum_{h1,h2} E_ab[h1] E_cd[h2] R_{h1,h2}
//     = sum_{k,j} Et_ab[k] Et_cd[j] (-1)^{|j|} R0_{k+j},
//   Et[k] = sum_{t>=k} E_t C(t,k) (-i q)^{t-k}   (per Cartesian axis),
// where R0 is the ordinary three-index Hermite Coulomb table evaluated at the
// complex displacement X = P'-Q' (R0^n_{t+1} = t R0^{n+1}_{t-1} + X R0^{n+1}_t)
// seeded with common * (-2 rho)^n F_n(T). The London phase therefore lives in
// complex one-dimensional Hermite coefficients and one R0 table serves every
// component quartet of a primitive quartet.
//
// Kernels.
//   seed_kernel:      one lane per primitive quartet: pair data
//                     (overlap.cpp:60-91), Boys argument and scaling choice
//                     (eri.cpp:162-210), complex Boys F_0..F_L, and the R0
//                     seeds s_n = pref * common * (-2 rho)^n F_n plus X.
//   transform_kernel: one 256-thread (or, for small items, 64-thread)
//                     workgroup per work item (shell quartet x
//                     contraction rows x ket-component range). Per primitive
//                     quartet the R0 table is built level by level in LDS,
//                     the ket transform W[k][cd] += sum_j Et_cd R0[k+j] is
//                     accumulated in registers over ket primitive pairs, and
//                     after each bra primitive pair the bra transform
//                     out[ab][cd] += sum_k Et_ab[k] W[k][cd] runs from LDS.
//   scatter kernels:  four symmetry images of canonical blocks
//                     (eri.cpp:777-805) and the Cholesky diagonal extraction
//                     (cholesky.cpp:305-312).
#include "gh_boys.h"
#include "gh_hermite.h"
#include "gh_internal.h"

namespace gh {

__constant__ double c_gl16x[16];
__constant__ double c_gl16w[16];
__constant__ double c_gl32x[32];
__constant__ double c_gl32w[32];
// Cartesian components of l = 0..kMaxShellL, x-major (types.cpp:221-237),
// packed x | y << 8 | z << 16, shell l starting at nherm(l - 1).
__constant__ unsigned c_cart[nherm(kMaxShellL)];

__device__ inline unsigned cart_of(int l, int c) { return c_cart[l * (l + 1) * (l + 2) / 6 + c]; }

// R0 entries of degree <= L sorted by degree (t | u << 6 | v << 12 | lexicographic index << 18),
// for L = 0..kColumnMaxL, the list of L at offset C(L + 3, 4).
constexpr int kColMaxL = kColumnMaxL;
constexpr int kColListSize = (kColMaxL + 1) * (kColMaxL + 2) * (kColMaxL + 3) * (kColMaxL + 4) / 24;
__constant__ unsigned c_rlist_all[kColListSize];
__device__ __host__ inline int rlist_offset(int L) { return L * (L + 1) * (L + 2) * (L + 3) / 24; }
// The R0 recurrence step of each entry of those lists (same offsets):
// x = lex | src1 << 16, y = src2 | axis << 16 | coef << 18, where the entry is
// X[axis] * R[src1] + coef * R[src2] of the level above (axis 3: the seed).
__constant__ int2 c_rstep_all[kColListSize];
// Lexicographic offsets lex_off(L, t, u) for L, t, u <= kColumnMaxL.
constexpr int kLexDim = kColMaxL + 1;
__constant__ int c_lexoff_all[kLexDim * kLexDim * kLexDim];
__device__ inline const int* lexoff_table(int L) { return c_lexoff_all + L * kLexDim * kLexDim; }

// ---------------------------------------------------------------------------
// Pair data (overlap.cpp:60-91, same operation order)
// ---------------------------------------------------------------------------
struct PairD {
    double p;
    double P[3];
    double q[3];
    zc Pc[3];
    double K;
    zc london;
};

__device__ inline void make_pair(double ea, const double* A, const double* ka, double eb, const double* B,
                                 const double* kb, PairD& o) {
    o.p = ea + eb;
    const double mu = ea * eb / o.p;
#pragma unroll
    for (int i = 0; i < 3; ++i) o.P[i] = (ea * A[i] + eb * B[i]) / o.p;
#pragma unroll
    for (int i = 0; i < 3; ++i) o.q[i] = kb[i] - ka[i];
    const double dx = A[0] - B[0], dy = A[1] - B[1], dz = A[2] - B[2];
    o.K = exp(-mu * (dx * dx + dy * dy + dz * dz));
    const double phase = -(o.q[0] * o.P[0] + o.q[1] * o.P[1] + o.q[2] * o.P[2]);
    const double damping = -(o.q[0] * o.q[0] + o.q[1] * o.q[1] + o.q[2] * o.q[2]) / (4.0 * o.p);
    o.london = zexp(zc{damping, phase});
#pragma unroll
    for (int i = 0; i < 3; ++i) o.Pc[i] = zc{o.P[i], -0.5 * o.q[i] / o.p};
}

__device__ inline void record_error(DevError* err, int code, long long item, zc z, int order) {
    if (atomicCAS(&err->code, 0, code) == 0) {
        err->item = item;
        err->zr = z.x;
        err->zi = z.y;
        err->order = order;
    }
}

// ---------------------------------------------------------------------------
// Seed kernels
//
// Phase A (seed_kernel): one lane per primitive quartet. Arguments in the
// asymptotic region (and domain errors) are finished here; arguments that
// need the adaptive quadrature are appended to one of 11 lists by their
// starting segment count, power-series arguments to a twelfth
// (workgroup-aggregated atomics).
// Phase B (seed_quad_kernel): persistent wavefronts take 64 list entries at
// a time, the series bin first, then the quadrature bins most expensive
// first, so a wavefront's lanes run arguments of
// similar length. Each argument is evaluated on its own, so the results do
// not depend on the order of the lists.
//
// Scratch layout (ints): bin b's count at b * kSeedCounterStride and the
// work counter at kBoysBins * kSeedCounterStride (each on a cache line of its
// own: every workgroup of phase A adds to them), then from kSeedScratchHead
// bin b's primitive-quartet indices at kSeedScratchHead + b * total, the
// quartet index of each entry at kSeedScratchHead + (kBoysBins + b) * total,
// and the block map (the quartet holding the first primitive quartet of each
// workgroup of phase A, then the last quartet) at
// kSeedScratchHead + 2 * kBoysBins * total.
// ---------------------------------------------------------------------------
__device__ inline int boys_bin(zc z) {
    const double d = fmax(fabs(z.x), fabs(z.y));
    int s = (int)ceil(d / 8.0);  // the adaptive quadrature's starting segment count
    if (s < 1) s = 1;
    if (s <= 8) return s - 1;
    if (s <= 10) return 8;
    if (s <= 14) return 9;
    return 10;
}
// The power-series arguments (|z| <= 0.75) also go to phase B, in a bin of
// their own: their per-order series cost each wavefront holding one of them
// far more than the asymptotic form costs the others.
constexpr int kSeriesBin = 11;
static_assert(kBoysBins == 12, "boys_bin covers 11 quadrature bins and the series bin");

// Asymptotic form and domain error, as boys_eval decides them; false when the
// argument needs the power series or the quadrature (phase B).
template <int NMAX>
__device__ inline bool boys_direct(zc z, int nmax, bool scaled, zc (&F)[NMAX + 1], int& st) {
    const double mag = zabs(z);
    double err = 0.0;
    if (mag <= 0.75) return false;
    const double asym = 2.0 * (double)nmax + 60.0;
    if (z.x >= asym && fabs(z.y) <= 0.5 * z.x) {
        st = boys_asymptotic<NMAX>(z, nmax, scaled, F, err);
        return true;
    }
    if (mag > 160.0) {
        st = kErrBoysDomain;
        return true;
    }
    return false;  // needs the adaptive quadrature
}

// Seeds of one primitive quartet. Phase A (QUAD false) returns false and sets
// bin when the argument needs the quadrature; phase B (QUAD true) evaluates it.
template <int NMAX, bool QUAD>
__device__ inline bool seed_unit(const SeedQuartet& Q, int qidx, long long local, const DevShell* __restrict__ shells,
                                 const double* __restrict__ exps, const double* __restrict__ kappa,
                                 zc* __restrict__ seeds, DevError* err, const double* __restrict__ radial, int& bin) {
    const int ab = (int)(local / Q.kcd), cd = (int)(local % Q.kcd);
    const DevShell& A = shells[Q.sh[0]];
    const DevShell& B = shells[Q.sh[1]];
    const DevShell& C = shells[Q.sh[2]];
    const DevShell& D = shells[Q.sh[3]];
    const int pa = ab / B.nprim, pb = ab % B.nprim, pc = cd / D.nprim, pd = cd % D.nprim;
    PairD P1, P2;
    make_pair(exps[A.prim_off + pa], A.center, kappa + 3 * Q.sh[0], exps[B.prim_off + pb], B.center,
              kappa + 3 * Q.sh[1], P1);
    make_pair(exps[C.prim_off + pc], C.center, kappa + 3 * Q.sh[2], exps[D.prim_off + pd], D.center,
              kappa + 3 * Q.sh[3], P2);
    const double p = P1.p, q = P2.p;
    const double rho = p * q / (p + q);
    const zc X[3] = {zsub(P1.Pc[0], P2.Pc[0]), zsub(P1.Pc[1], P2.Pc[1]), zsub(P1.Pc[2], P2.Pc[2])};
    const zc T = zscale(rho, zadd(zadd(zmul(X[0], X[0]), zmul(X[1], X[1])), zmul(X[2], X[2])));
    const bool scaled = T.x < 0.0;
    zc F[NMAX + 1];
    const int L = Q.L;
    int st = kErrNone;
    if constexpr (QUAD) {
        int region, segs;
        double berr;
        st = boys_eval<NMAX>(T, L, scaled, F, region, berr, segs);
    } else {
        if (!boys_direct<NMAX>(T, L, scaled, F, st)) {
            bin = zabs(T) <= 0.75 ? kSeriesBin : boys_bin(T);
            return false;
        }
    }
    zc* out = seeds + Q.seed_off + local * (L + 4);
    if (st != kErrNone) {
        record_error(err, st, qidx, T, L);
#pragma unroll
        for (int n = 0; n <= NMAX; ++n) F[n] = zc{0.0, 0.0};
    }
    zc common = zmul(P1.london, P2.london);
    if (scaled) {
        double rd2 = 0.0, tw2 = 0.0, twc = 0.0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double rd = P1.P[i] - P2.P[i];
            const double tw = P1.q[i] + P2.q[i];
            const double wc = (p * P1.P[i] + q * P2.P[i]) / (p + q);
            rd2 += rd * rd;
            tw2 += tw * tw;
            twc += tw * wc;
        }
        common = zexp(zc{-rho * rd2 - tw2 / (4.0 * (p + q)), -twc});
    }
    double pref = 2.0 * pow(kPi, 2.5) / (p * q * sqrt(p + q)) * P1.K * P2.K;
    if (radial)
        pref *= radial[A.prim_off + pa] * radial[B.prim_off + pb] * radial[C.prim_off + pc] *
                radial[D.prim_off + pd];
    out[0] = X[0];
    out[1] = X[1];
    out[2] = X[2];
    const zc base = zscale(pref, common);
    double rf = 1.0;
#pragma unroll
    for (int n = 0; n <= NMAX; ++n) {
        if (n <= L) out[3 + n] = zmul(zscale(rf, base), F[n]);
        rf *= -2.0 * rho;
    }
    return true;
}

__device__ inline int find_quartet(const SeedQuartet* __restrict__ quartets, int nq, long long gid) {
    int lo = 0, hi = nq - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (quartets[mid].pq_begin <= gid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Quartet index of the primitive quartet gid among quartets [lo, hi].
__device__ inline int find_quartet_in(const SeedQuartet* __restrict__ quartets, int lo, int hi, long long gid) {
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (quartets[mid].pq_begin <= gid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

constexpr int kSeedBlock = 256;  // lanes per workgroup of phase A
constexpr int kSeedCounterStride = 32;
static_assert((kBoysBins + 1) * kSeedCounterStride <= kSeedScratchHead, "seed counters fit the scratch head");

// Block map of phase A: the quartet of each block's first primitive quartet,
// and (entry nblocks) the last quartet, so that a lane searches only the few
// quartets its block spans instead of the whole list.
__global__ void seed_blockmap_kernel(const SeedQuartet* __restrict__ quartets, int nq, long long total, int nblocks,
                                     int* __restrict__ blkq) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nblocks) blkq[b] = find_quartet(quartets, nq, (long long)b * kSeedBlock);
    else if (b == nblocks) blkq[b] = nq - 1;
}

template <int NMAX>
__global__ __launch_bounds__(kSeedBlock) void seed_kernel(const SeedQuartet* __restrict__ quartets, int nq,
                                                          long long total, const DevShell* __restrict__ shells,
                                                          const double* __restrict__ exps,
                                                          const double* __restrict__ kappa, zc* __restrict__ seeds,
                                                          DevError* err, const double* __restrict__ radial,
                                                          int* __restrict__ scratch) {
    __shared__ int s_cnt[kBoysBins];   // the workgroup's count per bin
    __shared__ int s_base[kBoysBins];  // the workgroup's first position per bin
    const int tid = threadIdx.x;
    if (tid < kBoysBins) s_cnt[tid] = 0;
    __syncthreads();
    const long long gid = (long long)blockIdx.x * blockDim.x + tid;
    const int* blkq = scratch + kSeedScratchHead + 2 * (long long)kBoysBins * total;
    int bin = -1, qi = 0;
    if (gid < total) {
        qi = find_quartet_in(quartets, blkq[blockIdx.x], blkq[blockIdx.x + 1], gid);
        const SeedQuartet Q = quartets[qi];
        seed_unit<NMAX, false>(Q, qi, gid - Q.pq_begin, shells, exps, kappa, seeds, err, radial, bin);
    }
    // append the deferred arguments: positions within the workgroup from LDS
    // atomics (one per wavefront and bin), then one global atomic per bin
    const int lane = tid % 64;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    int my_off = 0;
#pragma unroll
    for (int bb = 0; bb < kBoysBins; ++bb) {
        const unsigned long long mask = __ballot(bin == bb);
        if (mask == 0) continue;
        const int leader = __ffsll((long long)mask) - 1;
        int off = 0;
        if (lane == leader) off = atomicAdd(&s_cnt[bb], __popcll(mask));
        off = __shfl(off, leader);
        if (bin == bb) my_off = off + __popcll(mask & below);
    }
    __syncthreads();
    if (tid < kBoysBins && s_cnt[tid] > 0) s_base[tid] = atomicAdd(scratch + tid * kSeedCounterStride, s_cnt[tid]);
    __syncthreads();
    if (bin >= 0) {
        const long long at = s_base[bin] + my_off;
        scratch[kSeedScratchHead + (long long)bin * total + at] = (int)gid;
        scratch[kSeedScratchHead + (long long)(kBoysBins + bin) * total + at] = qi;
    }
}

template <int NMAX>
__global__ __launch_bounds__(64) void seed_quad_kernel(const SeedQuartet* __restrict__ quartets, int nq,
                                                       long long total, const DevShell* __restrict__ shells,
                                                       const double* __restrict__ exps,
                                                       const double* __restrict__ kappa, zc* __restrict__ seeds,
                                                       DevError* err, const double* __restrict__ radial,
                                                       int* __restrict__ scratch) {
    int counts[kBoysBins];
    int all = 0;
#pragma unroll
    for (int bb = 0; bb < kBoysBins; ++bb) {
        counts[bb] = scratch[bb * kSeedCounterStride];
        all += counts[bb];
    }
    const int lane = threadIdx.x;
    for (;;) {
        int start = 0;
        if (lane == 0) start = atomicAdd(scratch + kBoysBins * kSeedCounterStride, 64);
        start = __shfl(start, 0);
        if (start >= all) break;  // every wavefront reaches this once the lists are taken
        const int idx = start + lane;
        int off = 0, bin = -1, pos = 0;
#pragma unroll
        for (int bb = kBoysBins - 1; bb >= 0; --bb) {  // series, then most expensive bin first
            if (bin < 0 && idx < off + counts[bb]) {
                bin = bb;
                pos = idx - off;
            }
            off += counts[bb];
        }
        if (bin >= 0) {
            const long long gid = scratch[kSeedScratchHead + (long long)bin * total + pos];
            const int qi = scratch[kSeedScratchHead + (long long)(kBoysBins + bin) * total + pos];
            const SeedQuartet Q = quartets[qi];
            int unused;
            seed_unit<NMAX, true>(Q, qi, gid - Q.pq_begin, shells, exps, kappa, seeds, err, radial, unused);
        }
    }
}

// Boys probe: one lane per argument (giaohip_boys).
__global__ __launch_bounds__(64) void boys_probe_kernel(const zc* __restrict__ z, long long n, int nmax, int scaled,
                                                        zc* __restrict__ out, int* __restrict__ region,
                                                        double* __restrict__ errest, int* __restrict__ segments,
                                                        DevError* err) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    zc F[kMaxBoysOrder + 1];
    int reg = -1, segs = 0;
    double e = 0.0;
    const int st = boys_eval<kMaxBoysOrder>(z[i], nmax, scaled != 0, F, reg, e, segs);
    if (st != kErrNone) record_error(err, st, i, z[i], nmax);
#pragma unroll
    for (int k = 0; k <= kMaxBoysOrder; ++k)
        if (k <= nmax) out[i * (nmax + 1) + k] = st == kErrNone ? F[k] : zc{0.0, 0.0};
    region[i] = reg;
    errest[i] = e;
    segments[i] = segs;
}

// ---------------------------------------------------------------------------
// Transform kernel
// ---------------------------------------------------------------------------
constexpr int kNT = 256;
constexpr int kWaveNT = 64;  // the one-wavefront form
constexpr int kMaxW = 8;  // W entries per thread
constexpr int kMaxO = 8;  // output entries per thread

// threadIdx.x, opaque to the optimizer: what is derived from it is computed
// where it is used rather than hoisted out of loops into live registers.
__device__ inline int opaque_tid(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

struct LdsLayout {
    int offL, listL, offAB, tuvAB;  // int offsets (in ints)
    int skey, sdec;                 // ket slots of a pass: sort keys, decode (int2)
    int seeds, eket, ebra, work;    // zc offsets (in zc) from zc base
    int kfac, bfac;                 // contraction x normalization factors (doubles)
    int zc_base;                    // byte offset of the zc region
};

__device__ __host__ inline int align16(int bytes) { return (bytes + 15) & ~15; }

__device__ __host__ inline LdsLayout lds_layout(int la, int lb, int lc, int ld, int tile, int* total_bytes) {
    const int L = la + lb + lc + ld, Lab = la + lb, Lcd = lc + ld;
    LdsLayout s;
    int ints = 0;
    s.offL = ints;
    ints += (L + 1) * (L + 1);
    s.listL = ints;
    ints += nherm(L);
    s.offAB = ints;
    ints += (Lab + 1) * (Lab + 1);
    s.tuvAB = ints;
    ints += nherm(Lab);
    s.skey = ints;
    ints += tile;
    ints = (ints + 1) & ~1;
    s.sdec = ints;
    ints += 2 * tile;
    s.zc_base = align16(ints * 4);
    int z = 0;
    s.seeds = z;
    z += L + 4;
    s.eket = z;
    z += 3 * (lc + 1) * (ld + 1) * (Lcd + 1);
    s.ebra = z;
    z += 3 * (la + 1) * (lb + 1) * (Lab + 1);
    s.kfac = z;
    z += (tile + 1) / 2;
    s.bfac = z;
    z += (ncart(la) * ncart(lb) + 1) / 2;
    s.work = z;
    const int r0 = 2 * nherm(L), wz = nherm(Lab) * tile;
    z += r0 > wz ? r0 : wz;
    if (total_bytes) *total_bytes = s.zc_base + z * 16;
    return s;
}

// Lexicographic (t, u, v) layout of all Hermite indices of degree <= D:
// index(t,u,v) = off(t,u) + v with off(t,u) = NH(D) - NH(D-t) + u(D-t+1) - u(u-1)/2
// (entries with first index >= t form NH(D-t); within t, u' < u contribute
// D-t-u'+1 each).
__device__ __host__ inline int lex_off(int D, int t, int u) {
    return nherm(D) - nherm(D - t) + u * (D - t + 1) - u * (u - 1) / 2;
}

// One work item of the generic transform. Every thread of the NT-thread
// workgroup calls it with the same item; smem_raw is the dynamic LDS region.
// NT = 256 (four wavefronts) or 64 (one wavefront: items whose table entries
// and transform rows fit one wavefront's registers; a CU then holds four
// times as many items, which hides the latency of the per-primitive-quartet
// steps, and the workgroup barriers are a single wavefront's).
template <int NT>
__device__ __forceinline__ void transform_item(const ClassArgs ca, const WorkItem it,
                                               const DevShell* __restrict__ shells, const double* __restrict__ exps,
                                               const double* __restrict__ coefs, const double* __restrict__ norms,
                                               const double* __restrict__ kappa, const zc* __restrict__ seeds,
                                               zc* __restrict__ out, unsigned char* smem_raw) {
    // the item's own class (mixed-class launches) or the launch's
    const int cls = it.cls;
    const int la = cls ? (cls & 15) : ca.l[0], lb = cls ? ((cls >> 4) & 15) : ca.l[1];
    const int lc = cls ? ((cls >> 8) & 15) : ca.l[2], ld = cls ? ((cls >> 12) & 15) : ca.l[3];
    const int L = la + lb + lc + ld, Lab = la + lb, Lcd = lc + ld;
    const int tile = cls ? (cls >> 16) : ca.tile;
    const LdsLayout lay = lds_layout(la, lb, lc, ld, tile, nullptr);
    int* ibase = reinterpret_cast<int*>(smem_raw);
    zc* zbase = reinterpret_cast<zc*>(smem_raw + lay.zc_base);
    int* offL = ibase + lay.offL;
    int* listL = ibase + lay.listL;
    int* offAB = ibase + lay.offAB;
    int* tuvAB = ibase + lay.tuvAB;
    int* skey = ibase + lay.skey;
    int2* sdec = reinterpret_cast<int2*>(ibase + lay.sdec);
    zc* sseed = zbase + lay.seeds;
    zc* eket = zbase + lay.eket;
    zc* ebra = zbase + lay.ebra;
    zc* work = zbase + lay.work;
    double* kfac = reinterpret_cast<double*>(zbase + lay.kfac);
    double* bfac = reinterpret_cast<double*>(zbase + lay.bfac);
    const int tid = threadIdx.x;

    // Index tables (parallel fills).
    for (int e = tid; e < (L + 1) * (L + 1); e += NT) {
        const int t = e / (L + 1), u = e % (L + 1);
        offL[e] = t + u <= L ? lex_off(L, t, u) : -1;
    }
    for (int e = tid; e < (Lab + 1) * (Lab + 1); e += NT) {
        const int t = e / (Lab + 1), u = e % (Lab + 1);
        if (t + u <= Lab) {
            const int o = lex_off(Lab, t, u);
            offAB[e] = o;
            for (int v = 0; v <= Lab - t - u; ++v) tuvAB[o + v] = t | (u << 6) | (v << 12);
        } else {
            offAB[e] = -1;
        }
    }
    // R0 entries sorted by degree: degree d, then t descending, then u descending.
    for (int e = tid; e < nherm(L); e += NT) {
        int d = 0;
        while (nherm(d) <= e) ++d;
        int r = e - (d > 0 ? nherm(d - 1) : 0);
        int t = d;
        while (r >= d - t + 1) {
            r -= d - t + 1;
            --t;
        }
        const int u = d - t - r, v = d - t - u;
        listL[e] = t | (u << 6) | (v << 12) | ((lex_off(L, t, u) + v) << 18);
    }
    __syncthreads();

    const DevShell A = shells[it.sh[0]];
    const DevShell B = shells[it.sh[1]];
    const DevShell C = shells[it.sh[2]];
    const DevShell D = shells[it.sh[3]];
    // centres read through global memory: a runtime axis index into a copied
    // struct would place the copies in scratch
    const double *Ac = shells[it.sh[0]].center, *Bc = shells[it.sh[1]].center;
    const double *Cc = shells[it.sh[2]].center, *Dc = shells[it.sh[3]].center;
    const int nca = ncart(la), ncb = ncart(lb), ncc = ncart(lc), ncd = ncart(ld);
    const int nab = nca * ncb;
    const int NHab = nherm(Lab);
    const int nsel = it.cd1 - it.cd0;
    const int Kb = B.nprim, Kd = D.nprim;
    const int Kab = A.nprim * Kb, Kcd = C.nprim * Kd;
    const int stride_pq = L + 4;
    const int ebra_i = (lb + 1) * (Lab + 1), ebra_a = (la + 1) * ebra_i;
    const int eket_i = (ld + 1) * (Lcd + 1), eket_a = (lc + 1) * eket_i;
    // e / NHab as a multiply-high: exact for every e < 2^32 / NHab
    const unsigned nh_inv = NHab > 1 ? 0xffffffffu / (unsigned)NHab + 1u : 0u;

    // The R0 recurrence step of entries tid and tid + NT (fixed for the
    // item): rs_a = lex | src1 << 16, rs_b = src2 | axis << 16 | (coef) << 18,
    // axis 3 marking the seed entry.
    int rs_a[2], rs_b[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        const int e = tid + r * NT;
        rs_a[r] = 0;
        rs_b[r] = 3 << 16;
        if (e < nherm(L)) {
            const int pk = listL[e];
            const int t = pk & 63, u = (pk >> 6) & 63, v = (pk >> 12) & 63, lex = pk >> 18;
            int s1 = 0, s2 = 0, ax = 3, cf = 0;
            if (t > 0) {
                ax = 0;
                s1 = offL[(t - 1) * (L + 1) + u] + v;
                if (t > 1) s2 = offL[(t - 2) * (L + 1) + u] + v, cf = t - 1;
            } else if (u > 0) {
                ax = 1;
                s1 = offL[u - 1] + v;
                if (u > 1) s2 = offL[u - 2] + v, cf = u - 1;
            } else if (v > 0) {
                ax = 2;
                s1 = v - 1;
                if (v > 1) s2 = v - 2, cf = v - 1;
            }
            rs_a[r] = lex | (s1 << 16);
            rs_b[r] = s2 | (ax << 16) | (cf << 18);
        }
    }

    // Seeds of the next primitive quartet (lanes below L + 4), loaded one
    // primitive quartet ahead so that their latency overlaps the current one;
    // the order is the loop order below (bra pair, ket pair), restarting for
    // each pass over the ket components.
    const long long npq = (long long)Kab * Kcd;
    const zc* seed_lane = seeds + it.seed_off + (tid < stride_pq ? tid : 0);
    zc seed_next = tid < stride_pq ? seed_lane[0] : zc{0.0, 0.0};

    for (int t0 = 0; t0 < nsel; t0 += tile) {
        const int tl = min(tile, nsel - t0);
        zc acc[kMaxO];
#pragma unroll
        for (int r = 0; r < kMaxO; ++r) acc[r] = zc{0.0, 0.0};
        // The pass's ket components in slot order: sorted by their loop
        // bounds (cx + dx, cy + dy, cz + dz), so that the lanes of a wavefront
        // mostly run the same trip counts. sdec[slot] = (the six Cartesian
        // exponents, 4 bits each; cc | dc << 8). (The previous pass's reads of
        // sdec all precede the barrier below.)
        for (int sl = tid; sl < tl; sl += NT) {
            const int cdcomp = it.cd0 + t0 + sl;
            const unsigned pcp = cart_of(lc, cdcomp / ncd), pdp = cart_of(ld, cdcomp % ncd);
            const int a = (pcp & 255) + (pdp & 255), b = ((pcp >> 8) & 255) + ((pdp >> 8) & 255);
            const int c = (pcp >> 16) + (pdp >> 16);
            skey[sl] = (((a << 10) | (b << 5) | c) << 12) | sl;
        }
        __syncthreads();
        for (int sl = tid; sl < tl; sl += NT) {
            const int key = skey[sl];
            int rank = 0;
            for (int s2 = 0; s2 < tl; ++s2) rank += skey[s2] < key ? 1 : 0;
            const int cdcomp = it.cd0 + t0 + sl;
            const int cc = cdcomp / ncd, dc = cdcomp % ncd;
            const unsigned pcp = cart_of(lc, cc), pdp = cart_of(ld, dc);
            sdec[rank] = make_int2((int)((pcp & 15) | ((pdp & 15) << 4) | (((pcp >> 8) & 15) << 8) |
                                         (((pdp >> 8) & 15) << 12) | ((pcp >> 16) << 16) | ((pdp >> 16) << 20)),
                                   cc | (dc << 8));
        }
        // (the first ket step's barrier orders these writes before their reads)

        for (int pab = 0; pab < Kab; ++pab) {
            const int pa = pab / Kb, pb = pab % Kb;
            const double ea = exps[A.prim_off + pa], eb = exps[B.prim_off + pb];
            const double p = ea + eb;
            // bra 1-D complex Hermite tables
            for (int e = tid; e < 3 * (la + 1) * (lb + 1); e += NT) {
                const int axis = e / ((la + 1) * (lb + 1));
                const int i = (e / (lb + 1)) % (la + 1), j = e % (lb + 1);
                const double Pc = (ea * Ac[axis] + eb * Bc[axis]) / p;
                const double qv = kappa[3 * it.sh[1] + axis] - kappa[3 * it.sh[0] + axis];
                build_phase_hermite(i, j, Pc - Ac[axis], Pc - Bc[axis], p, qv,
                                    ebra + axis * ebra_a + i * ebra_i + j * (Lab + 1));
            }
            // contraction x normalization of every bra component pair, the
            // product formed in the order ((cA nA) cB) nB
            for (int ab = tid; ab < nab; ab += NT) {
                const int cac = ab / ncb, cbc = ab % ncb;
                bfac[ab] = coefs[A.coef_off + it.x[0] * A.nprim + pa] * norms[A.norm_off + cac * A.nprim + pa] *
                           coefs[B.coef_off + it.x[1] * B.nprim + pb] * norms[B.norm_off + cbc * B.nprim + pb];
            }
            zc w[kMaxW];
#pragma unroll
            for (int r = 0; r < kMaxW; ++r) w[r] = zc{0.0, 0.0};

            for (int pcd = 0; pcd < Kcd; ++pcd) {
                const int pc = pcd / Kd, pd = pcd % Kd;
                const double ec = exps[C.prim_off + pc], ed = exps[D.prim_off + pd];
                const double q = ec + ed;
                __syncthreads();  // previous primitive quartet finished with LDS
                if (tid < stride_pq) {
                    sseed[tid] = seed_next;
                    long long nx = (long long)pab * Kcd + pcd + 1;
                    if (nx == npq) nx = 0;
                    seed_next = seed_lane[nx * stride_pq];
                }
                for (int e = tid; e < 3 * (lc + 1) * (ld + 1); e += NT) {
                    const int axis = e / ((lc + 1) * (ld + 1));
                    const int i = (e / (ld + 1)) % (lc + 1), j = e % (ld + 1);
                    const double Qc = (ec * Cc[axis] + ed * Dc[axis]) / q;
                    const double qv = kappa[3 * it.sh[3] + axis] - kappa[3 * it.sh[2] + axis];
                    zc* dst = eket + axis * eket_a + i * eket_i + j * (Lcd + 1);
                    build_phase_hermite(i, j, Qc - Cc[axis], Qc - Dc[axis], q, qv, dst);
                    // the ket side's (-1)^(tau + phi + chi), one axis at a time
                    for (int h = 1; h <= i + j; h += 2) dst[h] = zneg(dst[h]);
                }
                for (int sl = tid; sl < tl; sl += NT) {
                    const int cc = sdec[sl].y & 255, dc = sdec[sl].y >> 8;
                    kfac[sl] = coefs[C.coef_off + it.x[2] * C.nprim + pc] * norms[C.norm_off + cc * C.nprim + pc] *
                               coefs[D.coef_off + it.x[3] * D.nprim + pd] * norms[D.norm_off + dc * D.nprim + pd];
                }
                __syncthreads();
                // R0 table: level L holds s_L; level n from level n+1.
                zc* buf0 = work;
                zc* buf1 = work + nherm(L);
                if (tid == 0) buf0[0] = sseed[3 + L];
                const zc X0 = sseed[0], X1 = sseed[1], X2 = sseed[2];
                for (int n = L - 1; n >= 0; --n) {
                    __syncthreads();
                    const zc* cur = ((L - n) & 1) ? buf0 : buf1;
                    zc* nxt = ((L - n) & 1) ? buf1 : buf0;
                    const int cnt = nherm(L - n);
#pragma unroll
                    for (int r = 0; r < 2; ++r) {
                        if (tid + r * NT < cnt) {
                            const int lex = rs_a[r] & 0xffff, s1 = rs_a[r] >> 16;
                            const int s2 = rs_b[r] & 0xffff, ax = (rs_b[r] >> 16) & 3, cf = rs_b[r] >> 18;
                            zc val;
                            if (ax == 3) {
                                val = sseed[3 + n];
                            } else {
                                val = zmul(ax == 0 ? X0 : (ax == 1 ? X1 : X2), cur[s1]);
                                if (cf > 0) {
                                    const zc r2 = cur[s2];
                                    val.x += (double)cf * r2.x;
                                    val.y += (double)cf * r2.y;
                                }
                            }
                            nxt[lex] = val;
                        }
                    }
                    for (int e = tid + 2 * NT; e < cnt; e += NT) {
                        const int pk = listL[e];
                        const int t = pk & 63, u = (pk >> 6) & 63, v = (pk >> 12) & 63, lex = pk >> 18;
                        zc val;
                        if (t > 0) {
                            const int i1 = offL[(t - 1) * (L + 1) + u] + v;
                            val = zmul(X0, cur[i1]);
                            if (t > 1) {
                                const zc r2 = cur[offL[(t - 2) * (L + 1) + u] + v];
                                val.x += (double)(t - 1) * r2.x;
                                val.y += (double)(t - 1) * r2.y;
                            }
                        } else if (u > 0) {
                            val = zmul(X1, cur[offL[u - 1] + v]);
                            if (u > 1) {
                                const zc r2 = cur[offL[u - 2] + v];
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
                __syncthreads();
                const zc* R0 = (L & 1) ? buf1 : buf0;
                // Ket transform into registers. The entry decode is redone per
                // primitive quartet (opaque_tid): hoisted out of the loop it
                // would hold kMaxW copies of it in registers.
                const int nW = NHab * tl;
                const int tid_k = opaque_tid(tid);
#pragma unroll
                for (int r = 0; r < kMaxW; ++r) {
                    const int e = tid_k + r * NT;
                    if (e < nW) {
                        const int cdl = NHab > 1 ? (int)__umulhi((unsigned)e, nh_inv) : e;
                        const int k = e - cdl * NHab;
                        const int pk = tuvAB[k];
                        const int t = pk & 63, u = (pk >> 6) & 63, v = (pk >> 12) & 63;
                        const int dx6 = sdec[cdl].x;
                        const int cx = dx6 & 15, dx = (dx6 >> 4) & 15, cy = (dx6 >> 8) & 15;
                        const int dy = (dx6 >> 12) & 15, cz = (dx6 >> 16) & 15, dz = dx6 >> 20;
                        const zc* Ex = eket + 0 * eket_a + cx * eket_i + dx * (Lcd + 1);
                        const zc* Ey = eket + 1 * eket_a + cy * eket_i + dy * (Lcd + 1);
                        const zc* Ez = eket + 2 * eket_a + cz * eket_i + dz * (Lcd + 1);
                        // sum_{tau,phi,chi} (-1)^(tau+phi+chi) Ex Ey Ez R0, one axis at a time
                        zc s{0.0, 0.0};
                        for (int tau = 0; tau <= cx + dx; ++tau) {
                            zc st{0.0, 0.0};
                            for (int phi = 0; phi <= cy + dy; ++phi) {
                                const zc* rr = R0 + offL[(t + tau) * (L + 1) + (u + phi)] + v;
                                zc stp{0.0, 0.0};
                                for (int chi = 0; chi <= cz + dz; ++chi) zmac(stp, Ez[chi], rr[chi]);
                                zmac(st, Ey[phi], stp);
                            }
                            zmac(s, Ex[tau], st);
                        }
                        const double sc = kfac[cdl];
                        w[r].x = fma(sc, s.x, w[r].x);
                        w[r].y = fma(sc, s.y, w[r].y);
                    }
                }
            }
            __syncthreads();
            // W to LDS: work[k * tl + cdl]
            {
                const int nW = NHab * tl;
#pragma unroll
                for (int r = 0; r < kMaxW; ++r) {
                    const int e = tid + r * NT;
                    if (e < nW) {
                        const int cdl = NHab > 1 ? (int)__umulhi((unsigned)e, nh_inv) : e;
                        work[(e - cdl * NHab) * tl + cdl] = w[r];
                    }
                }
            }
            __syncthreads();
            // Bra transform.
            const int nO = nab * tl;
            const int tid_b = opaque_tid(tid);
#pragma unroll
            for (int r = 0; r < kMaxO; ++r) {
                const int o = tid_b + r * NT;
                if (o < nO) {
                    const int abcomp = o / tl, cdl = o % tl;
                    const int cac = abcomp / ncb, cbc = abcomp % ncb;
                    const unsigned pap = cart_of(la, cac), pbp = cart_of(lb, cbc);
                    const int ax = pap & 255, ay = (pap >> 8) & 255, az = pap >> 16;
                    const int bx = pbp & 255, by = (pbp >> 8) & 255, bz = pbp >> 16;
                    const zc* Ex = ebra + 0 * ebra_a + ax * ebra_i + bx * (Lab + 1);
                    const zc* Ey = ebra + 1 * ebra_a + ay * ebra_i + by * (Lab + 1);
                    const zc* Ez = ebra + 2 * ebra_a + az * ebra_i + bz * (Lab + 1);
                    // sum_{t,u,v} Ex Ey Ez W, one axis at a time
                    zc s{0.0, 0.0};
                    for (int t = 0; t <= ax + bx; ++t) {
                        zc st{0.0, 0.0};
                        for (int u = 0; u <= ay + by; ++u) {
                            const zc* ww = work + (offAB[t * (Lab + 1) + u]) * tl + cdl;
                            zc stu{0.0, 0.0};
                            for (int v = 0; v <= az + bz; ++v) zmac(stu, Ez[v], ww[v * tl]);
                            zmac(st, Ey[u], stu);
                        }
                        zmac(s, Ex[t], st);
                    }
                    const double sc = bfac[abcomp];
                    acc[r].x = fma(sc, s.x, acc[r].x);
                    acc[r].y = fma(sc, s.y, acc[r].y);
                }
            }
            __syncthreads();
        }
        // Write the tile.
        const int nO = nab * tl;
#pragma unroll
        for (int r = 0; r < kMaxO; ++r) {
            const int o = tid + r * NT;
            if (o < nO) {
                const int abcomp = o / tl, cdl = o % tl;
                const int cac = abcomp / ncb, cbc = abcomp % ncb;
                const int cc = sdec[cdl].y & 255, dc = sdec[cdl].y >> 8;
                const long long ia = (long long)it.x[0] * nca + cac, ib = (long long)it.x[1] * ncb + cbc;
                const long long ic = (long long)it.x[2] * ncc + cc, id = (long long)it.x[3] * ncd + dc;
                out[it.out_base + ia * it.stride[0] + ib * it.stride[1] + ic * it.stride[2] + id * it.stride[3]] =
                    acc[r];
            }
        }
    }
}

// One workgroup per work item.
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void transform_kernel(
    ClassArgs ca, const WorkItem* __restrict__ items, const DevShell* __restrict__ shells,
    const double* __restrict__ exps, const double* __restrict__ coefs, const double* __restrict__ norms,
    const double* __restrict__ kappa, const zc* __restrict__ seeds, zc* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    transform_item<NT>(ca, items[blockIdx.x], shells, exps, coefs, norms, kappa, seeds, out, smem_raw);
}

// Persistent form: the grid's workgroups take items from a device work queue
// (`queue` counts the items handed out; it is zero at launch) in list order,
// which the host sorts longest first, until the list is exhausted. Every
// workgroup leaves the loop once the counter has passed nitems.
__global__ __launch_bounds__(kNT) __attribute__((amdgpu_waves_per_eu(4))) void transform_queue_kernel(
    ClassArgs ca, const WorkItem* __restrict__ items, int nitems, int* __restrict__ queue,
    const DevShell* __restrict__ shells, const double* __restrict__ exps, const double* __restrict__ coefs,
    const double* __restrict__ norms, const double* __restrict__ kappa, const zc* __restrict__ seeds,
    zc* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(queue, 1);
        __syncthreads();
        const int i = next;
        if (i >= nitems) break;
        transform_item<kNT>(ca, items[i], shells, exps, coefs, norms, kappa, seeds, out, smem_raw);
        __syncthreads();  // LDS and `next` are free for the next item
    }
}

__global__ void queue_reset_kernel(int* __restrict__ queue, int n) {
    if ((int)threadIdx.x < n) queue[threadIdx.x] = 0;
}

__global__ void zero_ints_kernel(int* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_zero_ints(int* p, int n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    zero_ints_kernel<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(p, n);
    return hipGetLastError();
}

// Four symmetry images of canonical blocks into an n^4 tensor (eri.cpp:777-805).
__global__ void tensor_scatter_kernel(const long long* __restrict__ block_off, const int* __restrict__ block_q,
                                      int nblocks, long long total, const DevShell* __restrict__ shells,
                                      const long long* __restrict__ ao_off, long long nao,
                                      const zc* __restrict__ blocks, zc* __restrict__ tensor) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    int lo = 0, hi = nblocks - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (block_off[mid] <= gid) lo = mid;
        else hi = mid - 1;
    }
    const int* q = block_q + 4 * lo;
    const long long na = shells[q[0]].ao_count, nb = shells[q[1]].ao_count;
    const long long nc = shells[q[2]].ao_count, nd = shells[q[3]].ao_count;
    long long k = gid - block_off[lo];
    const long long id = k % nd;
    k /= nd;
    const long long ic = k % nc;
    k /= nc;
    const long long ib = k % nb;
    const long long ia = k / nb;
    const zc v = blocks[gid];
    const zc vc{v.x, -v.y};
    const long long ga = ao_off[q[0]] + ia, gb = ao_off[q[1]] + ib, gc = ao_off[q[2]] + ic, gd = ao_off[q[3]] + id;
    const long long n1 = nao, n2 = nao * nao, n3 = n2 * nao;
    tensor[ga * n3 + gb * n2 + gc * n1 + gd] = v;
    tensor[gc * n3 + gd * n2 + ga * n1 + gb] = v;
    tensor[gb * n3 + ga * n2 + gd * n1 + gc] = vc;
    tensor[gd * n3 + gc * n2 + gb * n1 + ga] = vc;
}

// Diagonal D[nu,mu] = D[mu,nu] = Re (mu nu|nu mu) from (ab|ba) blocks
// (cholesky.cpp:305-312).
__global__ void diag_extract_kernel(const long long* __restrict__ block_off, const int* __restrict__ pairs,
                                    int npairs, const DevShell* __restrict__ shells,
                                    const long long* __restrict__ ao_off, long long nao,
                                    const zc* __restrict__ blocks, double* __restrict__ diag) {
    const int pr = blockIdx.x;
    if (pr >= npairs) return;
    const int a = pairs[2 * pr], b = pairs[2 * pr + 1];
    const int na = shells[a].ao_count, nb = shells[b].ao_count;
    const zc* blk = blocks + block_off[pr];
    for (int e = threadIdx.x; e < na * nb; e += blockDim.x) {
        const int i = e / nb, j = e % nb;
        const double v = blk[(((long long)i * nb + j) * nb + j) * na + i].x;
        const long long mu = ao_off[a] + i, nu = ao_off[b] + j;
        diag[nu * nao + mu] = v;
        if (a != b) diag[mu * nao + nu] = v;
    }
}

// Real zero-field columns from complex columns over all shell pairs
// (cholesky.cpp:250-258,271-279).
__global__ void real_columns_kernel(const zc* __restrict__ cols, const int* __restrict__ owner, long long nao,
                                    long long npiv, double* __restrict__ out) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n2 = nao * nao;
    if (gid >= npiv * n2) return;
    const long long cand = gid / n2, rest = gid % n2;
    const long long mu = rest / nao, nu = rest % nao;
    const int a = owner[mu], b = owner[nu];
    const zc* col = cols + cand * n2;
    double v;
    if (a < b) v = col[mu * nao + nu].x;
    else if (a > b) v = col[nu * nao + mu].x;
    else v = (col[mu * nao + nu].x + col[nu * nao + mu].x) * 0.5;
    out[gid] = v;
}

// ---------------------------------------------------------------------------
// Host-side launch wrappers
// ---------------------------------------------------------------------------
int transform_lds_bytes(const ClassArgs& ca) {
    int bytes = 0;
    lds_layout(ca.l[0], ca.l[1], ca.l[2], ca.l[3], ca.tile, &bytes);
    return bytes;
}

int transform_max_w(int nt) { return kMaxW * nt; }
int transform_max_o(int nt) { return kMaxO * nt; }

hipError_t upload_constants() {
    // Gauss-Legendre nodes by Newton iteration on P_n, the same procedure as
    // the reference's rule construction (boys.cpp:27-109).
    auto rule = [](int n, double* x, double* w) {
        for (int i = 0; i < n / 2; ++i) {
            double root = std::cos(kPi * ((double)i + 0.75) / ((double)n + 0.5));
            double derivative = 0.0;
            for (int it = 0; it < 20; ++it) {
                double prev = 1.0, cur = root;
                for (int deg = 2; deg <= n; ++deg) {
                    const double next = ((2.0 * (double)deg - 1.0) * root * cur - ((double)deg - 1.0) * prev) /
                                        (double)deg;
                    prev = cur;
                    cur = next;
                }
                derivative = (double)n * (root * cur - prev) / (root * root - 1.0);
                const double update = cur / derivative;
                root -= update;
                if (std::fabs(update) < 4.0 * 2.220446049250313e-16) break;
            }
            const double weight = 2.0 / ((1.0 - root * root) * derivative * derivative);
            x[i] = -root;
            x[n - 1 - i] = root;
            w[i] = weight;
            w[n - 1 - i] = weight;
        }
    };
    double x16[16], w16[16], x32[32], w32[32];
    rule(16, x16, w16);
    rule(32, x32, w32);
    hipError_t e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_gl16x), x16, sizeof x16)) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_gl16w), w16, sizeof w16)) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_gl32x), x32, sizeof x32)) != hipSuccess) return e;
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_gl32w), w32, sizeof w32)) != hipSuccess) return e;
    unsigned cart[nherm(kMaxShellL)];
    int k = 0;
    for (int l = 0; l <= kMaxShellL; ++l)
        for (int x = l; x >= 0; --x)
            for (int y = l - x; y >= 0; --y) cart[k++] = (unsigned)x | ((unsigned)y << 8) | ((unsigned)(l - x - y) << 16);
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_cart), cart, sizeof cart)) != hipSuccess) return e;
    std::vector<unsigned> rlist(kColListSize);
    for (int L = 0; L <= kColMaxL; ++L) {
        unsigned* dst = rlist.data() + rlist_offset(L);
        int e = 0;
        for (int d = 0; d <= L; ++d)
            for (int t = d; t >= 0; --t)
                for (int u = d - t; u >= 0; --u) {
                    const int v = d - t - u;
                    dst[e++] = (unsigned)t | ((unsigned)u << 6) | ((unsigned)v << 12) |
                               ((unsigned)(lex_off(L, t, u) + v) << 18);
                }
    }
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_rlist_all), rlist.data(), rlist.size() * sizeof(unsigned))) != hipSuccess)
        return e;
    std::vector<int2> rstep(kColListSize);
    for (int L = 0; L <= kColMaxL; ++L)
        for (int i = 0; i < nherm(L); ++i) {
            const unsigned pk = rlist[rlist_offset(L) + i];
            const int t = pk & 63, u = (pk >> 6) & 63, v = (pk >> 12) & 63, lex = (int)(pk >> 18);
            int s1 = 0, s2 = 0, ax = 3, cf = 0;
            if (t > 0) {
                ax = 0;
                s1 = lex_off(L, t - 1, u) + v;
                if (t > 1) s2 = lex_off(L, t - 2, u) + v, cf = t - 1;
            } else if (u > 0) {
                ax = 1;
                s1 = lex_off(L, 0, u - 1) + v;
                if (u > 1) s2 = lex_off(L, 0, u - 2) + v, cf = u - 1;
            } else if (v > 0) {
                ax = 2;
                s1 = v - 1;
                if (v > 1) s2 = v - 2, cf = v - 1;
            }
            rstep[rlist_offset(L) + i] = make_int2(lex | (s1 << 16), s2 | (ax << 16) | (cf << 18));
        }
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_rstep_all), rstep.data(), rstep.size() * sizeof(int2))) != hipSuccess)
        return e;
    std::vector<int> lexoff(kLexDim * kLexDim * kLexDim, -1);
    for (int L = 0; L <= kColMaxL; ++L)
        for (int t = 0; t <= L; ++t)
            for (int u = 0; t + u <= L; ++u) lexoff[(L * kLexDim + t) * kLexDim + u] = lex_off(L, t, u);
    if ((e = hipMemcpyToSymbol(HIP_SYMBOL(c_lexoff_all), lexoff.data(), lexoff.size() * sizeof(int))) != hipSuccess)
        return e;
    return hipSuccess;
}

size_t seed_scratch_ints(long long total) {
    const size_t nblocks = (size_t)((total + kSeedBlock - 1) / kSeedBlock);
    return kSeedScratchHead + 2 * (size_t)kBoysBins * (size_t)total + nblocks + 1;
}

hipError_t launch_seed(int nmax_bucket, const SeedQuartet* q, int nq, long long total, const DevShell* shells,
                       const double* exps, const double* kappa, zc* seeds, DevError* err, int* scratch,
                       hipStream_t s, const double* radial) {
    if (total == 0) return hipSuccess;
    // the quadrature list counters start at zero (a fill kernel, so the
    // sequence holds no memset node when a caller captures it)
    hipError_t e = launch_zero_ints(scratch, kSeedScratchHead, s);
    if (e != hipSuccess) return e;
    const dim3 block(kSeedBlock);
    const int nblocks = (int)((total + kSeedBlock - 1) / kSeedBlock);
    const dim3 grid((unsigned)nblocks);
    seed_blockmap_kernel<<<dim3((unsigned)((nblocks + 1 + 255) / 256)), dim3(256), 0, s>>>(
        q, nq, total, nblocks, scratch + kSeedScratchHead + 2 * (long long)kBoysBins * total);
    const long long qwaves = (total + 63) / 64;
    const dim3 qgrid((unsigned)(qwaves < 4096 ? qwaves : 4096));
#define GH_SEED(N)                                                                                              \
    seed_kernel<N><<<grid, block, 0, s>>>(q, nq, total, shells, exps, kappa, seeds, err, radial, scratch);     \
    seed_quad_kernel<N><<<qgrid, dim3(64), 0, s>>>(q, nq, total, shells, exps, kappa, seeds, err, radial, scratch); \
    break;
    switch (nmax_bucket) {
        case 4: GH_SEED(4)
        case 8: GH_SEED(8)
        case 12: GH_SEED(12)
        case 16: GH_SEED(16)
        case 24: GH_SEED(24)
        default: GH_SEED(32)
    }
#undef GH_SEED
    return hipGetLastError();
}

hipError_t launch_transform(const ClassArgs& ca, const WorkItem* items, int nitems, const DevShell* shells,
                            const double* exps, const double* coefs, const double* norms, const double* kappa,
                            const zc* seeds, zc* out, hipStream_t s) {
    if (nitems == 0) return hipSuccess;
    const int bytes = transform_lds_bytes(ca);
    if (bytes > 64 * 1024) {
        static bool raised = false;
        if (!raised) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&transform_kernel<kNT>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            if (e != hipSuccess) return e;
            raised = true;
        }
    }
    transform_kernel<kNT><<<dim3((unsigned)nitems), dim3(kNT), bytes, s>>>(ca, items, shells, exps, coefs, norms,
                                                                           kappa, seeds, out);
    return hipGetLastError();
}

hipError_t launch_transform_mixed(int nt, int lds_bytes, const WorkItem* items, int nitems, const DevShell* shells,
                                  const double* exps, const double* coefs, const double* norms, const double* kappa,
                                  const zc* seeds, zc* out, int* queue, hipStream_t s) {
    if (nitems == 0) return hipSuccess;
    if (lds_bytes > 160 * 1024 || (nt != kNT && nt != kWaveNT)) return hipErrorInvalidValue;
    const ClassArgs none{{0, 0, 0, 0}, 1};  // every item carries its class
    if (nt == kWaveNT) {
        if (lds_bytes > 64 * 1024) {
            static bool raised_w = false;
            if (!raised_w) {
                hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&transform_kernel<kWaveNT>),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
                if (e != hipSuccess) return e;
                raised_w = true;
            }
        }
        transform_kernel<kWaveNT><<<dim3((unsigned)nitems), dim3(kWaveNT), lds_bytes, s>>>(
            none, items, shells, exps, coefs, norms, kappa, seeds, out);
        return hipGetLastError();
    }
    if (lds_bytes > 64 * 1024) {
        static bool raised = false;
        if (!raised) {
            hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&transform_kernel<kNT>),
                                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
            if (e != hipSuccess) return e;
            raised = true;
        }
    }
    // the queue form also holds one int of static LDS
    if (queue && lds_bytes + 256 <= 160 * 1024) {
        hipError_t e = hipSuccess;
        if (lds_bytes > 64 * 1024) {
            static bool raised_q = false;
            if (!raised_q) {
                e = hipFuncSetAttribute(reinterpret_cast<const void*>(&transform_queue_kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 256);
                if (e != hipSuccess) return e;
                raised_q = true;
            }
        }
        int per_cu = 0, dev = 0, cus = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&transform_queue_kernel),
                                                         kNT, lds_bytes);
        if (e != hipSuccess) return e;
        if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
        if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
        const long long resident = (long long)(per_cu > 0 ? per_cu : 1) * (cus > 0 ? cus : 1);
        const unsigned grid = (unsigned)(resident < nitems ? resident : nitems);
        queue_reset_kernel<<<dim3(1), dim3(64), 0, s>>>(queue, 1);
        transform_queue_kernel<<<dim3(grid), dim3(kNT), lds_bytes, s>>>(none, items, nitems, queue, shells, exps,
                                                                         coefs, norms, kappa, seeds, out);
        return hipGetLastError();
    }
    transform_kernel<kNT><<<dim3((unsigned)nitems), dim3(kNT), lds_bytes, s>>>(none, items, shells, exps, coefs,
                                                                               norms, kappa, seeds, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Cholesky column kernel: one wavefront per bra shell pair of a column, for
// the pivot's single ket component (cholesky.cpp:203-282 evaluates the same
// elements). The ket side needs only that component's 1-D tables and the
// products over its box; R0 is built level by level in the wavefront's LDS;
// lanes take bra Hermite indices for V[k] = sum_j E_cd[j] (-1)^|j| R0[k+j],
// then bra components for out[ab] = sum_k E_ab[k] V[k]. Normalization and
// contraction coefficients enter as in transform_kernel.
// LDS (complex): 2 nh(L) + 3 (la+1)(lb+1)(Lab+1) + 3 (Lcd+1) + box + nh(Lab) + nab + L + 4.
// ---------------------------------------------------------------------------
__device__ inline void wave_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64) void chol_column_kernel(ClassArgs ca, const WorkItem* __restrict__ items,
                                                         const DevShell* __restrict__ shells,
                                                         const double* __restrict__ exps,
                                                         const double* __restrict__ coefs,
                                                         const double* __restrict__ norms,
                                                         const double* __restrict__ kappa,
                                                         const zc* __restrict__ seeds, zc* __restrict__ out,
                                                         int nitems, int per_pair_bytes) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // ca.tile carries P, the bra pairs per wavefront: one per group of
    // stride = 64 / P lanes, each with its own LDS region; the groups share
    // only the wave fences
    const int stride = 64 / ca.tile;
    const int sub = (int)threadIdx.x / stride;
    const int lane = (int)threadIdx.x % stride;
    const int item = blockIdx.x * ca.tile + sub;
    if (item >= nitems) return;  // no block barriers below: the rest of the wavefront continues
    unsigned char* smem_pair = smem_raw + (size_t)sub * per_pair_bytes;
    const int la = ca.l[0], lb = ca.l[1], lc = ca.l[2], ld = ca.l[3];
    const int L = la + lb + lc + ld, Lab = la + lb, Lcd = lc + ld;
    const int NHL = nherm(L), NHab = nherm(Lab);
    const WorkItem it = items[item];
    const DevShell A = shells[it.sh[0]];
    const DevShell B = shells[it.sh[1]];
    const DevShell C = shells[it.sh[2]];
    const DevShell D = shells[it.sh[3]];
    // centres read through global memory: a runtime axis index into a copied
    // struct would place the copies in scratch
    const double *Ac = shells[it.sh[0]].center, *Bc = shells[it.sh[1]].center;
    const double *Cc = shells[it.sh[2]].center, *Dc = shells[it.sh[3]].center;
    const int nca = ncart(la), ncb = ncart(lb), ncd = ncart(ld);
    const int nab = nca * ncb;
    const int cdcomp = it.cd0;
    const int cc = cdcomp / ncd, dc = cdcomp % ncd;
    const unsigned pcp = cart_of(lc, cc), pdp = cart_of(ld, dc);
    const int cpow[3] = {(int)(pcp & 255), (int)((pcp >> 8) & 255), (int)(pcp >> 16)};
    const int dpow[3] = {(int)(pdp & 255), (int)((pdp >> 8) & 255), (int)(pdp >> 16)};
    const int bx = cpow[0] + dpow[0] + 1, by = cpow[1] + dpow[1] + 1, bz = cpow[2] + dpow[2] + 1;
    const int nbox = bx * by * bz;
    const int ebra_i = (lb + 1) * (Lab + 1), ebra_a = (la + 1) * ebra_i;
    zc* buf0 = reinterpret_cast<zc*>(smem_pair);
    zc* buf1 = buf0 + NHL;
    zc* ebra = buf1 + NHL;
    zc* e1d = ebra + 3 * ebra_a;
    zc* eket = e1d + 3 * (Lcd + 1);
    zc* vv = eket + nbox;
    zc* acc = vv + NHab;
    zc* sseed = acc + nab;
    const int2* rs = c_rstep_all + rlist_offset(L);
    const int* lexL = lexoff_table(L);
    const int* lexAB = lexoff_table(Lab);
    for (int o = lane; o < nab; o += stride) acc[o] = zc{0.0, 0.0};
    const int Kb = B.nprim, Kd = D.nprim;
    const int Kab = A.nprim * Kb, Kcd = C.nprim * Kd;
    for (int pab = 0; pab < Kab; ++pab) {
        const int pa = pab / Kb, pb = pab % Kb;
        const double ea = exps[A.prim_off + pa], eb = exps[B.prim_off + pb];
        const double p = ea + eb;
        wave_barrier();
        for (int e = lane; e < 3 * (la + 1) * (lb + 1); e += stride) {
            const int axis = e / ((la + 1) * (lb + 1));
            const int i = (e / (lb + 1)) % (la + 1), j = e % (lb + 1);
            const double Pc = (ea * Ac[axis] + eb * Bc[axis]) / p;
            const double qv = kappa[3 * it.sh[1] + axis] - kappa[3 * it.sh[0] + axis];
            build_phase_hermite(i, j, Pc - Ac[axis], Pc - Bc[axis], p, qv,
                                ebra + axis * ebra_a + i * ebra_i + j * (Lab + 1));
        }
        for (int k = lane; k < NHab; k += stride) vv[k] = zc{0.0, 0.0};
        for (int pcd = 0; pcd < Kcd; ++pcd) {
            const int pc = pcd / Kd, pd = pcd % Kd;
            const double ec = exps[C.prim_off + pc], ed = exps[D.prim_off + pd];
            const double q = ec + ed;
            wave_barrier();  // the previous primitive quartet is done with the buffers
            const zc* sp = seeds + it.seed_off + (long long)(pab * Kcd + pcd) * (L + 4);
            if (lane < L + 4) sseed[lane] = sp[lane];
            if (lane < 3) {
                const int axis = lane;
                const double Qc = (ec * Cc[axis] + ed * Dc[axis]) / q;
                const double qv = kappa[3 * it.sh[3] + axis] - kappa[3 * it.sh[2] + axis];
                build_phase_hermite((int)((pcp >> (8 * axis)) & 255), (int)((pdp >> (8 * axis)) & 255), Qc - Cc[axis], Qc - Dc[axis], q, qv,
                                    e1d + axis * (Lcd + 1));
            }
            wave_barrier();
            for (int e = lane; e < nbox; e += stride) {
                const int jt = e / (by * bz), ju = (e / bz) % by, jv = e % bz;
                zc v = zmul(zmul(e1d[jt], e1d[(Lcd + 1) + ju]), e1d[2 * (Lcd + 1) + jv]);
                if ((jt + ju + jv) & 1) v = zneg(v);
                eket[e] = v;
            }
            // R0 table: level L holds s_L; level n from level n + 1.
            if (lane == 0) buf0[0] = sseed[3 + L];
            const zc X0 = sseed[0], X1 = sseed[1], X2 = sseed[2];
            for (int n = L - 1; n >= 0; --n) {
                wave_barrier();
                const zc* cur = ((L - n) & 1) ? buf0 : buf1;
                zc* nxt = ((L - n) & 1) ? buf1 : buf0;
                const int cnt = nherm(L - n);
                for (int e = lane; e < cnt; e += stride) {
                    const int2 st = rs[e];
                    const int lex = st.x & 0xffff, s1 = st.x >> 16;
                    const int s2 = st.y & 0xffff, ax = (st.y >> 16) & 3, cf = st.y >> 18;
                    zc val;
                    if (ax == 3) {
                        val = sseed[3 + n];
                    } else {
                        val = zmul(ax == 0 ? X0 : (ax == 1 ? X1 : X2), cur[s1]);
                        if (cf > 0) {
                            const zc r2 = cur[s2];
                            val.x += (double)cf * r2.x;
                            val.y += (double)cf * r2.y;
                        }
                    }
                    nxt[lex] = val;
                }
            }
            wave_barrier();
            const zc* R0 = (L & 1) ? buf1 : buf0;
            const double sc = coefs[C.coef_off + it.x[2] * C.nprim + pc] * norms[C.norm_off + cc * C.nprim + pc] *
                              coefs[D.coef_off + it.x[3] * D.nprim + pd] * norms[D.norm_off + dc * D.nprim + pd];
            for (int e = lane; e < NHab; e += stride) {
                const unsigned pk = c_rlist_all[rlist_offset(Lab) + e];
                const int kt = pk & 63, ku = (pk >> 6) & 63, kv = (pk >> 12) & 63, k = (int)(pk >> 18);
                zc s{0.0, 0.0};
                for (int jt = 0; jt < bx; ++jt)
                    for (int ju = 0; ju < by; ++ju) {
                        const zc* rr = R0 + lexL[(kt + jt) * kLexDim + ku + ju] + kv;
                        const zc* ee = eket + (jt * by + ju) * bz;
                        for (int jv = 0; jv < bz; ++jv) zmac(s, ee[jv], rr[jv]);
                    }
                vv[k].x = fma(sc, s.x, vv[k].x);
                vv[k].y = fma(sc, s.y, vv[k].y);
            }
        }
        wave_barrier();
        for (int o = lane; o < nab; o += stride) {
            const int cac = o / ncb, cbc = o % ncb;
            const unsigned pap = cart_of(la, cac), pbp = cart_of(lb, cbc);
            const int ax = pap & 255, ay = (pap >> 8) & 255, az = pap >> 16;
            const int bxp = pbp & 255, byp = (pbp >> 8) & 255, bzp = pbp >> 16;
            const zc* Ex = ebra + 0 * ebra_a + ax * ebra_i + bxp * (Lab + 1);
            const zc* Ey = ebra + 1 * ebra_a + ay * ebra_i + byp * (Lab + 1);
            const zc* Ez = ebra + 2 * ebra_a + az * ebra_i + bzp * (Lab + 1);
            // sum_{t,u,v} Ex Ey Ez V, one axis at a time
            zc s{0.0, 0.0};
            for (int t = 0; t <= ax + bxp; ++t) {
                zc st{0.0, 0.0};
                for (int u = 0; u <= ay + byp; ++u) {
                    const zc* ww = vv + lexAB[t * kLexDim + u];
                    zc stu{0.0, 0.0};
                    for (int v = 0; v <= az + bzp; ++v) zmac(stu, Ez[v], ww[v]);
                    zmac(st, Ey[u], stu);
                }
                zmac(s, Ex[t], st);
            }
            const double sc = coefs[A.coef_off + it.x[0] * A.nprim + pa] * norms[A.norm_off + cac * A.nprim + pa] *
                              coefs[B.coef_off + it.x[1] * B.nprim + pb] * norms[B.norm_off + cbc * B.nprim + pb];
            acc[o].x = fma(sc, s.x, acc[o].x);
            acc[o].y = fma(sc, s.y, acc[o].y);
        }
    }
    const long long ic = (long long)it.x[2] * ncart(lc) + cc, id = (long long)it.x[3] * ncd + dc;
    for (int o = lane; o < nab; o += stride) {
        const int cac = o / ncb, cbc = o % ncb;
        const long long ia = (long long)it.x[0] * nca + cac, ib = (long long)it.x[1] * ncb + cbc;
        out[it.out_base + ia * it.stride[0] + ib * it.stride[1] + ic * it.stride[2] + id * it.stride[3]] = acc[o];
    }
}

int chol_column_lds_bytes(int la, int lb, int lc, int ld, int comp) {
    const int L = la + lb + lc + ld, Lab = la + lb, Lcd = lc + ld;
    const int ncd = ncart(ld), cc = comp / ncd, dc = comp % ncd;
    int cpow[3], dpow[3];
    {
        int k = 0;
        for (int x = lc; x >= 0; --x)
            for (int y = lc - x; y >= 0; --y, ++k)
                if (k == cc) cpow[0] = x, cpow[1] = y, cpow[2] = lc - x - y;
        k = 0;
        for (int x = ld; x >= 0; --x)
            for (int y = ld - x; y >= 0; --y, ++k)
                if (k == dc) dpow[0] = x, dpow[1] = y, dpow[2] = ld - x - y;
    }
    const int nbox = (cpow[0] + dpow[0] + 1) * (cpow[1] + dpow[1] + 1) * (cpow[2] + dpow[2] + 1);
    const int zcs = 2 * nherm(L) + 3 * (la + 1) * (lb + 1) * (Lab + 1) + 3 * (Lcd + 1) + nbox + nherm(Lab) +
                    ncart(la) * ncart(lb) + L + 4;
    return zcs * (int)sizeof(zc);
}

hipError_t launch_chol_column(const ClassArgs& ca, int comp, const WorkItem* items, int nitems, const DevShell* shells,
                              const double* exps, const double* coefs, const double* norms, const double* kappa,
                              const zc* seeds, zc* out, hipStream_t s) {
    if (nitems == 0) return hipSuccess;
    const int L = ca.l[0] + ca.l[1] + ca.l[2] + ca.l[3];
    if (L > kColMaxL) return hipErrorInvalidValue;
    const int bytes = chol_column_lds_bytes(ca.l[0], ca.l[1], ca.l[2], ca.l[3], comp);
    if (bytes > 64 * 1024) return hipErrorInvalidValue;
    // several bra pairs per wavefront when the pairs are small (their
    // loops are narrower than a wavefront and each pair is mostly latency)
    const int lds_cap = 64 * 1024;
    const int P = (L <= 3 && 8 * bytes <= lds_cap)    ? 8
                  : (L <= 6 && 4 * bytes <= lds_cap)  ? 4
                  : (L <= 10 && 2 * bytes <= lds_cap) ? 2
                                                      : 1;
    ClassArgs cp = ca;
    cp.tile = P;
    chol_column_kernel<<<dim3((unsigned)((nitems + P - 1) / P)), dim3(64), P * bytes, s>>>(
        cp, items, shells, exps, coefs, norms, kappa, seeds, out, nitems, bytes);
    return hipGetLastError();
}

hipError_t launch_boys_probe(const zc* z, long long n, int nmax, int scaled, zc* out, int* region, double* errest,
                             int* segments, DevError* err, hipStream_t s) {
    if (n == 0) return hipSuccess;
    boys_probe_kernel<<<dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s>>>(z, n, nmax, scaled, out, region, errest,
                                                                         segments, err);
    return hipGetLastError();
}

hipError_t launch_tensor_scatter(const long long* block_off, const int* block_q, int nblocks, long long total,
                                 const DevShell* shells, const long long* ao_off, long long nao, const zc* blocks,
                                 zc* tensor, hipStream_t s) {
    if (total == 0) return hipSuccess;
    tensor_scatter_kernel<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>(
        block_off, block_q, nblocks, total, shells, ao_off, nao, blocks, tensor);
    return hipGetLastError();
}

hipError_t launch_diag_extract(const long long* block_off, const int* pairs, int npairs, const DevShell* shells,
                               const long long* ao_off, long long nao, const zc* blocks, double* diag,
                               hipStream_t s) {
    if (npairs == 0) return hipSuccess;
    diag_extract_kernel<<<dim3((unsigned)npairs), dim3(256), 0, s>>>(block_off, pairs, npairs, shells, ao_off, nao,
                                                                    blocks, diag);
    return hipGetLastError();
}

hipError_t launch_real_columns(const zc* cols, const int* owner, long long nao, long long npiv, double* out,
                               hipStream_t s) {
    const long long total = npiv * nao * nao;
    if (total == 0) return hipSuccess;
    real_columns_kernel<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>(cols, owner, nao, npiv, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Nuclear attraction (nuclear.cpp:24-207, compute_nuclear_attraction_matrix
// nuclear.cpp:282-312): V[mu,nu] = sum_C -Z_C <mu| 1/|r - C| |nu> with London
// orbitals. The reference builds a four-index auxiliary R^n_{tuv} with the
// pair's wave term -i q; as for the ERI bra pair, the phase is folded into
// the Hermite coefficients (E~, built by build_phase_hermite), leaving the
// ordinary three-index R0 at the complex displacement X = P' - C. One
// wavefront per ordered shell pair; lanes take component pairs for the
// contraction and (t,u,v) cells for the R0 levels.
constexpr int kNucSide = kNucMaxL + 1;
constexpr int kNucCube = kNucSide * kNucSide * kNucSide;

__global__ __launch_bounds__(64) void nuclear_kernel(int nshell, const DevShell* __restrict__ shells,
                                                     const double* __restrict__ exps,
                                                     const double* __restrict__ coefs,
                                                     const double* __restrict__ norms,
                                                     const double* __restrict__ kappa,
                                                     const double* __restrict__ nuclei, int nnuc,
                                                     const long long* __restrict__ ao_off, long long nao,
                                                     zc* __restrict__ out, DevError* err) {
    __shared__ zc R[2][kNucCube];
    __shared__ zc E[3][kNucSide * kNucSide * kNucSide];  // [axis][i][j][k], k <= i + j
    __shared__ zc seed[kNucSide];
    const int lane = threadIdx.x;
    const int sa = blockIdx.x / nshell, sb = blockIdx.x % nshell;
    const DevShell& A = shells[sa];
    const DevShell& B = shells[sb];
    const int la = A.l, lb = B.l, L = la + lb;
    const int na = A.ncart, nb = B.ncart, npair = na * nb;
    const int side = L + 1;
    for (int xa = 0; xa < A.ncontr; ++xa) {
        for (int xb = 0; xb < B.ncontr; ++xb) {
            zc acc[kNucMaxComp];
#pragma unroll
            for (int r = 0; r < kNucMaxComp; ++r) acc[r] = zc{0.0, 0.0};
            for (int pa = 0; pa < A.nprim; ++pa) {
                for (int pb = 0; pb < B.nprim; ++pb) {
                    const double ea = exps[A.prim_off + pa], eb = exps[B.prim_off + pb];
                    PairD P;
                    make_pair(ea, A.center, kappa + 3 * sa, eb, B.center, kappa + 3 * sb, P);
                    for (int e = lane; e < 3 * (la + 1) * (lb + 1); e += 64) {
                        const int axis = e / ((la + 1) * (lb + 1));
                        const int i = (e / (lb + 1)) % (la + 1), j = e % (lb + 1);
                        build_phase_hermite(i, j, P.P[axis] - A.center[axis], P.P[axis] - B.center[axis], P.p,
                                            P.q[axis], &E[axis][(i * kNucSide + j) * kNucSide]);
                    }
                    zc vsum[kNucMaxComp];
#pragma unroll
                    for (int r = 0; r < kNucMaxComp; ++r) vsum[r] = zc{0.0, 0.0};
                    for (int c = 0; c < nnuc; ++c) {
                        const double Z = nuclei[4 * c];
                        const double C[3] = {nuclei[4 * c + 1], nuclei[4 * c + 2], nuclei[4 * c + 3]};
                        const zc X[3] = {zsub(P.Pc[0], zc{C[0], 0.0}), zsub(P.Pc[1], zc{C[1], 0.0}),
                                         zsub(P.Pc[2], zc{C[2], 0.0})};
                        if (lane == 0) {
                            const zc T = zscale(P.p, zadd(zadd(zmul(X[0], X[0]), zmul(X[1], X[1])), zmul(X[2], X[2])));
                            const bool scaled = T.x < 0.0;
                            zc F[kNucMaxL + 1];
                            int region, segs;
                            double berr;
                            const int st = boys_eval<kNucMaxL>(T, L, scaled, F, region, berr, segs);
                            if (st != kErrNone) {
                                record_error(err, st, blockIdx.x, T, L);
#pragma unroll
                                for (int n = 0; n <= kNucMaxL; ++n) F[n] = zc{0.0, 0.0};
                            }
                            zc common = P.london;
                            if (scaled) {
                                double d2 = 0.0, qc = 0.0;
#pragma unroll
                                for (int i = 0; i < 3; ++i) {
                                    const double d = P.P[i] - C[i];
                                    d2 += d * d;
                                    qc += P.q[i] * C[i];
                                }
                                common = zexp(zc{-P.p * d2, -qc});
                            }
                            double rf = 1.0;
#pragma unroll
                            for (int n = 0; n <= kNucMaxL; ++n) {
                                if (n <= L) seed[n] = zscale(rf, zmul(common, F[n]));
                                rf *= -2.0 * P.p;
                            }
                        }
                        __syncthreads();
                        // R0 levels: level n from level n + 1, cells (t,u,v) with t+u+v <= L-n
                        int cur = 0;
                        if (lane == 0) R[0][0] = seed[L];
                        for (int n = L - 1; n >= 0; --n) {
                            __syncthreads();
                            const zc* rc = R[cur];
                            zc* rn = R[cur ^ 1];
                            const int deg = L - n;
                            for (int e = lane; e < side * side * side; e += 64) {
                                const int t = e / (side * side), u = (e / side) % side, v = e % side;
                                if (t + u + v > deg) continue;
                                zc val;
                                if (t > 0) {
                                    val = zmul(X[0], rc[((t - 1) * side + u) * side + v]);
                                    if (t > 1) val = zadd(val, zscale((double)(t - 1), rc[((t - 2) * side + u) * side + v]));
                                } else if (u > 0) {
                                    val = zmul(X[1], rc[(u - 1) * side + v]);
                                    if (u > 1) val = zadd(val, zscale((double)(u - 1), rc[(u - 2) * side + v]));
                                } else if (v > 0) {
                                    val = zmul(X[2], rc[v - 1]);
                                    if (v > 1) val = zadd(val, zscale((double)(v - 1), rc[v - 2]));
                                } else {
                                    val = seed[n];
                                }
                                rn[(t * side + u) * side + v] = val;
                            }
                            cur ^= 1;
                        }
                        __syncthreads();
                        const zc* R0 = R[cur];
#pragma unroll
                        for (int r = 0; r < kNucMaxComp; ++r) {
                            const int e = lane + 64 * r;
                            if (e < npair) {
                                const unsigned ca = cart_of(la, e / nb), cb = cart_of(lb, e % nb);
                                const int ax = ca & 255, ay = (ca >> 8) & 255, az = ca >> 16;
                                const int bx = cb & 255, by = (cb >> 8) & 255, bz = cb >> 16;
                                const zc* Ex = &E[0][(ax * kNucSide + bx) * kNucSide];
                                const zc* Ey = &E[1][(ay * kNucSide + by) * kNucSide];
                                const zc* Ez = &E[2][(az * kNucSide + bz) * kNucSide];
                                zc sx{0.0, 0.0};
                                for (int t = 0; t <= ax + bx; ++t) {
                                    zc sy{0.0, 0.0};
                                    for (int u = 0; u <= ay + by; ++u) {
                                        zc sz{0.0, 0.0};
                                        for (int v = 0; v <= az + bz; ++v)
                                            sz = zadd(sz, zmul(Ez[v], R0[(t * side + u) * side + v]));
                                        sy = zadd(sy, zmul(Ey[u], sz));
                                    }
                                    sx = zadd(sx, zmul(Ex[t], sy));
                                }
                                vsum[r] = zadd(vsum[r], zscale(-Z, sx));
                            }
                        }
                        __syncthreads();  // R, seed and E reused by the next nucleus / primitive pair
                    }
                    const double pref = 2.0 * kPi / P.p * P.K * coefs[A.coef_off + xa * A.nprim + pa] *
                                        coefs[B.coef_off + xb * B.nprim + pb];
#pragma unroll
                    for (int r = 0; r < kNucMaxComp; ++r) {
                        const int e = lane + 64 * r;
                        if (e < npair) {
                            const double nrm = norms[A.norm_off + (e / nb) * A.nprim + pa] *
                                               norms[B.norm_off + (e % nb) * B.nprim + pb];
                            acc[r] = zadd(acc[r], zscale(pref * nrm, vsum[r]));
                        }
                    }
                }
            }
#pragma unroll
            for (int r = 0; r < kNucMaxComp; ++r) {
                const int e = lane + 64 * r;
                if (e < npair) {
                    const long long mu = ao_off[sa] + (long long)xa * na + e / nb;
                    const long long nu = ao_off[sb] + (long long)xb * nb + e % nb;
                    out[mu * nao + nu] = acc[r];
                }
            }
        }
    }
}

hipError_t launch_nuclear(int nshell, const DevShell* shells, const double* exps, const double* coefs,
                          const double* norms, const double* kappa, const double* nuclei, int nnuc,
                          const long long* ao_off, long long nao, zc* out, DevError* err, hipStream_t s) {
    if (nshell == 0) return hipSuccess;
    nuclear_kernel<<<dim3((unsigned)nshell * (unsigned)nshell), dim3(64), 0, s>>>(
        nshell, shells, exps, coefs, norms, kappa, nuclei, nnuc, ao_off, nao, out, err);
    return hipGetLastError();
}

}  // namespace gh
