// gfx950 kernels of the device-resident pivoted Cholesky of the ordered-pair
// Coulomb matrix W[(mu,nu),(lam,sig)] = (nu mu|lam sig)
// (python/giao_integrals/cholesky.py:60-297, strict global pivoting).
//
// Pivots are taken in blocks. A block starts at the current maximum of the
// residual diagonal and names as candidates the B largest residuals in
// (value desc, index asc) order, which is np.argmax's order, so the first
// candidate is the pivot itself:
//   topk_kernel     the B candidates, from the per-block maxima of the last
//                   residual update (only the blocks holding a top-B maximum
//                   can hold a top-B entry);
//   coef_kernel     c[k][m] = conj(L[k][p_m]) for the factors k < rank4
//                   (rank rounded down to a multiple of 4);
//   project_mfma_kernel  proj[m] = sum_{k < rank4} c[k][m] L[k] for all
//                   candidates in one pass over the factors on the FP64
//                   matrix cores (the GEMV of cholesky.py:268 for B columns at
//                   once: B times fewer factor reads).
// Every accepted pivot then needs its raw ERI column (expand_kernel writes
// its seed quartets and work items from a static bra table and the pivot's
// ket (cholesky.cpp:203-282), the seed and column kernels of gh_eri.hip
// evaluate it) and accept_kernel: L[rank] = (column - proj) / sqrt(res[p]),
// res -= |L[rank]|^2 with the PSD check, clamp at 0, res[p] = 0 and per-block
// (max, first index) partials. The argmax kernels finish the first maximum.
// When that maximum is a candidate not yet taken, the next pivot needs no
// pass over all factors: catchup_mfma_kernel adds to its projection the terms
// of the factors from its projection's rank4 up to the current rank (the
// candidates never taken never get them); otherwise a new block starts.
//
// Every projection accumulates conj(L[k][p]) L[k] from zero in groups of four
// factor rows aligned at row 0, each group by the same matrix-core
// instruction sequence (a candidate's value does not depend on the other
// rows of the tile), and is subtracted from the raw column once. Blocked and
// one-at-a-time runs therefore produce bitwise identical factors and pivots
// for every block size. (chol_project_kernel, FMAs in increasing k, serves
// the Coulomb contraction of the finished factors.)
#include "gh_internal.h"

namespace gh {

// ---------------------------------------------------------------------------
// Column work expansion
// ---------------------------------------------------------------------------
__global__ void chol_expand_kernel(const CholUnit* __restrict__ units, int nunits, CholKet ket,
                                   SeedQuartet* __restrict__ sq, WorkItem* __restrict__ wi) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nunits) return;
    const CholUnit U = units[u];
    SeedQuartet q{};
    q.sh[0] = U.sh_a;
    q.sh[1] = U.sh_b;
    q.sh[2] = ket.sh_c;
    q.sh[3] = ket.sh_d;
    q.L = U.lab + ket.lcd;
    q.kab = U.kab;
    q.kcd = ket.kcd;
    q.pq_begin = U.p1 * ket.kcd;
    q.seed_off = (U.p2 + U.p1 * (long long)(ket.lcd + 4)) * ket.kcd;
    sq[u] = q;
    WorkItem w{};
    w.sh[0] = U.sh_a;
    w.sh[1] = U.sh_b;
    w.sh[2] = ket.sh_c;
    w.sh[3] = ket.sh_d;
    w.x[0] = U.xa;
    w.x[1] = U.xb;
    w.x[2] = ket.xc;
    w.x[3] = ket.xd;
    w.cd0 = ket.comp;
    w.cd1 = ket.comp + 1;
    w.seed_off = q.seed_off;
    w.out_base = U.out_base;
    w.stride[0] = ket.n;
    w.stride[1] = 1;
    w.stride[2] = 0;
    w.stride[3] = 0;
    wi[u] = w;
}

// ---------------------------------------------------------------------------
// Factor update
// ---------------------------------------------------------------------------
constexpr int kUpdNT = 256;
constexpr int kUpdPerThread = 2;

__device__ inline const zc* factor_row(const zc* const* chunks, int rows_per_chunk, long long n2, int k) {
    return chunks[k / rows_per_chunk] + (long long)(k % rows_per_chunk) * n2;
}

// Block reduction of (value, index): larger value wins, lower index on ties.
__device__ inline void argmax_combine(double& v, long long& i, double v2, long long i2) {
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

template <int NT>
__device__ inline void block_argmax(double v, long long i, double* sv, long long* si, double* out_v, long long* out_i) {
    const int tid = threadIdx.x;
    sv[tid] = v;
    si[tid] = i;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (tid < s) {
            double a = sv[tid];
            long long ia = si[tid];
            argmax_combine(a, ia, sv[tid + s], si[tid + s]);
            sv[tid] = a;
            si[tid] = ia;
        }
        __syncthreads();
    }
    if (tid == 0) {
        *out_v = sv[0];
        *out_i = si[0];
    }
}

// acc += c * v with a fixed operation order (shared by every projection).
__device__ __forceinline__ void cmac(zc& acc, const zc c, const zc v) {
    acc.x = fma(c.x, v.x, acc.x);
    acc.x = fma(-c.y, v.y, acc.x);
    acc.y = fma(c.x, v.y, acc.y);
    acc.y = fma(c.y, v.x, acc.y);
}

constexpr int kPartElems = kUpdNT * kUpdPerThread;  // residual entries per partial maximum
constexpr int kTopNT = 1024;

// (v, i) follows (pv, pi) in (value desc, index asc) order
__device__ inline bool follows(double v, long long i, double pv, long long pi) {
    return v < pv || (v == pv && i > pi);
}

// First maximum over the workgroup; every thread receives it.
template <int NT>
__device__ inline void block_first_max(double& v, long long& i, double* sv, long long* si) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double v2 = __shfl_xor(v, off);
        const long long i2 = __shfl_xor(i, off);
        argmax_combine(v, i, v2, i2);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        sv[wid] = v;
        si[wid] = i;
    }
    __syncthreads();
    v = sv[0];
    i = si[0];
    for (int w = 1; w < NT / 64; ++w) argmax_combine(v, i, sv[w], si[w]);
}

// The nb (<= kCholMaxBlock) largest positive residuals in (value desc, index
// asc) order. An entry among the top nb lies in a partial block whose
// (max, first index) is among the top nb partials: each partial ranked above
// it holds a distinct larger entry.
__global__ __launch_bounds__(kTopNT) void chol_topk_kernel(const double* __restrict__ res, long long n2,
                                                           const double* __restrict__ pmax,
                                                           const long long* __restrict__ pidx, int nparts, int nb,
                                                           long long* __restrict__ cand, int* __restrict__ ncand) {
    __shared__ double sv[kTopNT / 64];
    __shared__ long long si[kTopNT / 64];
    __shared__ long long blocks[kCholMaxBlock];
    double pv = INFINITY;
    long long pi = -1;
    int nblk = 0;
    for (int t = 0; t < nb; ++t) {
        double v = -1.0;
        long long i = 0x7fffffffffffffffLL;
        for (int b = threadIdx.x; b < nparts; b += kTopNT) {
            const double bv = pmax[b];
            const long long bi = pidx[b];
            if (follows(bv, bi, pv, pi)) argmax_combine(v, i, bv, bi);
        }
        block_first_max<kTopNT>(v, i, sv, si);
        if (!(v > 0.0)) break;
        if (threadIdx.x == 0) blocks[t] = i / kPartElems;
        pv = v;
        pi = i;
        nblk = t + 1;
    }
    __syncthreads();
    constexpr int kPer = kCholMaxBlock * kPartElems / kTopNT;
    double ev[kPer];
    long long ei[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int slot = e * kTopNT + threadIdx.x;
        const int blk = slot / kPartElems;
        ev[e] = -1.0;
        ei[e] = 0x7fffffffffffffffLL;
        if (blk < nblk) {
            const long long j = blocks[blk] * kPartElems + slot % kPartElems;
            if (j < n2) {
                ev[e] = res[j];
                ei[e] = j;
            }
        }
    }
    pv = INFINITY;
    pi = -1;
    int count = 0;
    for (int t = 0; t < nb; ++t) {
        double v = -1.0;
        long long i = 0x7fffffffffffffffLL;
#pragma unroll
        for (int e = 0; e < kPer; ++e)
            if (follows(ev[e], ei[e], pv, pi)) argmax_combine(v, i, ev[e], ei[e]);
        block_first_max<kTopNT>(v, i, sv, si);
        if (!(v > 0.0)) break;
        if (threadIdx.x == 0) cand[t] = i;
        pv = v;
        pi = i;
        count = t + 1;
    }
    if (threadIdx.x == 0) *ncand = count;
}

// coef[k][m] = conj(L[k][piv[m]]) for k < rank, m < nnew; zero for the
// padding slots nnew <= m < nbt.
__global__ void chol_coef_kernel(const zc* const* __restrict__ chunks, int rows_per_chunk, long long n2, int rank,
                                 const long long* __restrict__ piv, int nnew, int nbt, zc* __restrict__ coef) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)rank * nbt) return;
    const int k = (int)(t / nbt), m = (int)(t % nbt);
    zc v{0.0, 0.0};
    if (m < nnew) {
        const zc x = factor_row(chunks, rows_per_chunk, n2, k)[piv[m]];
        v = zc{x.x, -x.y};
    }
    coef[t] = v;
}

// Load through a global-address-space pointer (the factor rows are reached
// through a table of chunk pointers, which the compiler cannot place).
typedef __attribute__((address_space(1))) const double global_double;
__device__ __forceinline__ zc gload(const zc* p) {
    const global_double* g = (const global_double*)p;
    return zc{g[0], g[1]};
}

// proj[slots[m]][j] = sum_{k < rank} coef[k][m] L[k][j] for the m < nnew new
// candidates (NB >= nnew coefficient columns);
// NE entries per thread. The factor rows are read once for all slots, U rows
// per group with the next group's rows in flight during the current FMAs.
// Lanes past n2 read entry n2 - 1 and store nothing.
template <int NB, int NE, int U>
__global__ __launch_bounds__(256) void chol_project_kernel(const zc* const* __restrict__ chunks, int rows_per_chunk,
                                                           long long n2, int rank, const zc* __restrict__ coef,
                                                           int nnew, const int* __restrict__ slots,
                                                           zc* __restrict__ proj) {
    const long long base = (long long)blockIdx.x * 256 * NE + threadIdx.x;
    long long jc[NE];
    zc acc[NB][NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const long long j = base + e * 256;
        jc[e] = j < n2 ? j : n2 - 1;
#pragma unroll
        for (int m = 0; m < NB; ++m) acc[m][e] = zc{0.0, 0.0};
    }
    auto step = [&](const zc* ck, const zc (&v)[NE]) {
#pragma unroll
        for (int m = 0; m < NB; ++m) {
            const zc c = ck[m];
#pragma unroll
            for (int e = 0; e < NE; ++e) cmac(acc[m][e], c, v[e]);
        }
    };
    for (int k0 = 0; k0 < rank; k0 += rows_per_chunk) {
        const zc* chunk = chunks[k0 / rows_per_chunk];
        const int rows = min(rows_per_chunk, rank - k0);
        const int full = rows - rows % (2 * U);  // rows taken 2U at a time
        const zc* ck = coef + (long long)k0 * NB;
        // rows r..r+U-1 (clamped to the chunk: the loads past the last group
        // repeat row rows-1 and are not used)
        auto load = [&](zc (&v)[U][NE], int r) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int rr = min(r + u, rows - 1);
#pragma unroll
                for (int e = 0; e < NE; ++e) v[u][e] = gload(chunk + (long long)rr * n2 + jc[e]);
            }
        };
        zc a[U][NE], b[U][NE];
        load(a, 0);
        // the scheduling barriers keep each group's loads ahead of the
        // previous group's FMAs (the scheduler would otherwise sink them)
        for (int r = 0; r < full; r += 2 * U) {
            load(b, r + U);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < U; ++u) step(ck + (long long)(r + u) * NB, a[u]);
            __builtin_amdgcn_sched_barrier(0);
            load(a, r + 2 * U);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < U; ++u) step(ck + (long long)(r + U + u) * NB, b[u]);
            __builtin_amdgcn_sched_barrier(0);
        }
        for (int r = full; r < rows; ++r) {
            zc x[NE];
#pragma unroll
            for (int e = 0; e < NE; ++e) x[e] = gload(chunk + (long long)r * n2 + jc[e]);
            step(ck + (long long)r * NB, x);
        }
    }
#pragma unroll
    for (int m = 0; m < NB; ++m)
        if (m < nnew)
#pragma unroll
            for (int e = 0; e < NE; ++e)
                if (base + e * 256 < n2) proj[(long long)slots[m] * n2 + base + e * 256] = acc[m][e];
}

constexpr int kStage = 64;  // factor rows per coefficient stage in LDS

// ---------------------------------------------------------------------------
// Projections on the FP64 matrix cores. v_mfma_f64_16x16x4f64 computes
// D = A B + C for a 16x4 A, a 4x16 B and a 16x16 C; lane l supplies A[l%16][l/16]
// and B[l/16][l%16] and holds D[l/16 + 4 r][l%16] in register r.
// A = coefficients (candidate x factor row), B = factor rows (row x entry):
// per group of four factor rows and per complex part, two MFMAs,
//   re = A_re B_re + (-A_im) B_im + re,   im = A_re B_im + A_im B_re + im,
// in that order. Every projection (block start, catch-up) accumulates its
// factor rows in these groups of four, aligned at row 0, from zero, so a
// value never depends on which kernel or which block summed which group.
// ---------------------------------------------------------------------------
typedef double mf4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void mfma_group(mf4& re, mf4& im, double are, double aim, double anim, double bre,
                                           double bim) {
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bre, re, 0, 0, 0);
    re = __builtin_amdgcn_mfma_f64_16x16x4f64(anim, bim, re, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(are, bim, im, 0, 0, 0);
    im = __builtin_amdgcn_mfma_f64_16x16x4f64(aim, bre, im, 0, 0, 0);
}

// proj[slots[m]][j] = sum_{k < rows} coef[k][m] L[k][j] for m < nnew; rows is
// a multiple of 4. MT tiles of 16 candidates (coef row stride 16 MT); each
// wave takes NJ tiles of 16 entries; the workgroup stages kStage coefficient
// rows in LDS. The next group's factor loads are issued before the current
// group's MFMAs.
template <int MT, int NJ>
__global__ __launch_bounds__(256) void chol_project_mfma_kernel(const zc* const* __restrict__ chunks,
                                                                int rows_per_chunk, long long n2, int rows,
                                                                const zc* __restrict__ coef, int nnew,
                                                                const int* __restrict__ slots,
                                                                zc* __restrict__ proj) {
    constexpr int NB = 16 * MT;
    __shared__ zc sc[kStage * NB];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int li = lane % 16, lk = lane / 16;
    const long long j0 = ((long long)blockIdx.x * 4 + wave) * (16 * NJ);
    mf4 acc[MT][2][NJ];
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < NJ; ++t) acc[mt][0][t] = acc[mt][1][t] = mf4{0.0, 0.0, 0.0, 0.0};
    long long jj[NJ];
#pragma unroll
    for (int t = 0; t < NJ; ++t) {
        const long long j = j0 + 16 * t + li;
        jj[t] = j < n2 ? j : n2 - 1;  // lanes past n2 read a valid entry and store nothing
    }
    auto load = [&](zc (&b)[NJ], int k) {
        const zc* row = chunks[k / rows_per_chunk] + (long long)(k % rows_per_chunk) * n2;
#pragma unroll
        for (int t = 0; t < NJ; ++t) b[t] = gload(row + jj[t]);
    };
    for (int s0 = 0; s0 < rows; s0 += kStage) {
        const int srows = min(kStage, rows - s0);
        __syncthreads();  // the previous stage's coefficients are no longer read
        for (int i = threadIdx.x; i < srows * NB; i += 256) sc[i] = coef[(long long)s0 * NB + i];
        __syncthreads();
        zc b[NJ], bn[NJ];
        load(b, s0 + lk);
        for (int r = 0; r < srows; r += 4) {
            load(bn, s0 + min(r + 4, srows - 4) + lk);  // the next group (the last group reloads itself)
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                const zc a = sc[(r + lk) * NB + 16 * mt + li];
#pragma unroll
                for (int t = 0; t < NJ; ++t) mfma_group(acc[mt][0][t], acc[mt][1][t], a.x, a.y, -a.y, b[t].x, b[t].y);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int t = 0; t < NJ; ++t) b[t] = bn[t];
        }
    }
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int t = 0; t < NJ; ++t) {
            const long long j = j0 + 16 * t + li;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = 16 * mt + lk + 4 * q;
                if (m < nnew && j < n2) proj[(long long)slots[m] * n2 + j] = zc{acc[mt][0][t][q], acc[mt][1][t][q]};
            }
        }
}

// proj += sum_{from <= k < to} conj(L[k][pm]) L[k] in the groups of four rows
// of chol_project_mfma_kernel (from is a multiple of 4; rows past `to` of the
// last group are zero): the candidate is row 0 of the A tile, the other rows
// zero. Each wave takes 16 NJ entries.
constexpr int kCatchNJ = 4;
__global__ __launch_bounds__(256) void chol_catchup_mfma_kernel(const zc* const* __restrict__ chunks,
                                                                int rows_per_chunk, long long n2, int from, int to,
                                                                long long pm, zc* __restrict__ proj) {
    constexpr int NJ = kCatchNJ;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int li = lane % 16, lk = lane / 16;
    const long long j0 = ((long long)blockIdx.x * 4 + wave) * (16 * NJ);
    mf4 re[NJ], im[NJ];
    long long jj[NJ];
#pragma unroll
    for (int t = 0; t < NJ; ++t) {
        const long long j = j0 + 16 * t + li;
        jj[t] = j < n2 ? j : n2 - 1;
        const zc p0 = lk == 0 ? proj[jj[t]] : zc{0.0, 0.0};  // D row 0 = lanes 0..15, register 0
        re[t] = mf4{p0.x, 0.0, 0.0, 0.0};
        im[t] = mf4{p0.y, 0.0, 0.0, 0.0};
    }
    for (int k4 = from; k4 < to; k4 += 4) {
        const int k = k4 + lk;
        double are = 0.0, aim = 0.0;
        zc b[NJ];
        if (k < to) {
            const zc* row = chunks[k / rows_per_chunk] + (long long)(k % rows_per_chunk) * n2;
            if (li == 0) {
                const zc x = row[pm];
                are = x.x;
                aim = -x.y;
            }
#pragma unroll
            for (int t = 0; t < NJ; ++t) b[t] = gload(row + jj[t]);
        } else {
#pragma unroll
            for (int t = 0; t < NJ; ++t) b[t] = zc{0.0, 0.0};
        }
#pragma unroll
        for (int t = 0; t < NJ; ++t) mfma_group(re[t], im[t], are, aim, -aim, b[t].x, b[t].y);
    }
    if (lk == 0)
#pragma unroll
        for (int t = 0; t < NJ; ++t) {
            const long long j = j0 + 16 * t + li;
            if (j < n2) proj[j] = zc{re[t][0], im[t][0]};
        }
}

// Accepts the pivot st->pivot, whose raw column is `column` and whose
// projection onto the earlier factors is `pproj`: writes the new factor row,
// updates the residual diagonal (PSD check, clamp, res[p] = 0) and the
// per-block (max, first index) partials.
// With `real` set (zero-field or one-centre bases, real arithmetic in
// cholesky.py:175-176,269-271) the projected column is first symmetrized:
// (c[mu,nu] + c[nu,mu]) / 2, imaginary part dropped; n is the AO count.
__device__ inline zc chol_column_value(const zc* __restrict__ column, const zc* __restrict__ pproj, long long j,
                                       int real, long long n) {
    const zc a = column[j], q = pproj[j];
    zc v{a.x - q.x, a.y - q.y};
    if (real) {
        const long long jt = (j % n) * n + j / n;
        const zc at = column[jt], qt = pproj[jt];
        v = zc{(v.x + (at.x - qt.x)) * 0.5, 0.0};
    }
    return v;
}

__global__ __launch_bounds__(kUpdNT) void chol_accept_kernel(const zc* __restrict__ column,
                                                             const zc* __restrict__ pproj, int real, long long n,
                                                             long long n2, const CholState* __restrict__ st,
                                                             double* __restrict__ res, zc* __restrict__ out_row,
                                                             double psd_tol, int* __restrict__ flags,
                                                             double* __restrict__ pmax, long long* __restrict__ pidx) {
    __shared__ double sv[kUpdNT];
    __shared__ long long si[kUpdNT];
    const int tid = threadIdx.x;
    const long long base = (long long)blockIdx.x * kPartElems;
    const double root = sqrt(st->value);
    const long long p = st->pivot;
    long long j[kUpdPerThread];
    zc val[kUpdPerThread];
    double best = -1.0;
    long long bidx = 0x7fffffffffffffffLL;
#pragma unroll
    for (int e = 0; e < kUpdPerThread; ++e) {
        j[e] = base + e * kUpdNT + tid;
        val[e] = zc{0.0, 0.0};
        if (j[e] < n2) {
            const zc cv = chol_column_value(column, pproj, j[e], real, n);
            val[e] = zc{cv.x / root, cv.y / root};
            out_row[j[e]] = val[e];
            double nr = res[j[e]] - (val[e].x * val[e].x + val[e].y * val[e].y);
            if (nr < -psd_tol) atomicOr(flags, 1);
            nr = nr > 0.0 ? nr : 0.0;
            if (j[e] == p) nr = 0.0;
            res[j[e]] = nr;
            argmax_combine(best, bidx, nr, j[e]);
        }
    }
    block_argmax<kUpdNT>(best, bidx, sv, si, pmax + blockIdx.x, pidx + blockIdx.x);
}

__global__ __launch_bounds__(kUpdNT) void chol_argmax_partial_kernel(const double* __restrict__ res, long long n2,
                                                                     double* __restrict__ pmax,
                                                                     long long* __restrict__ pidx) {
    __shared__ double sv[kUpdNT];
    __shared__ long long si[kUpdNT];
    const long long base = (long long)blockIdx.x * kUpdNT * kUpdPerThread;
    double best = -1.0;
    long long bidx = 0x7fffffffffffffffLL;
    for (int e = 0; e < kUpdPerThread; ++e) {
        const long long j = base + e * kUpdNT + threadIdx.x;
        if (j < n2) argmax_combine(best, bidx, res[j], j);
    }
    block_argmax<kUpdNT>(best, bidx, sv, si, pmax + blockIdx.x, pidx + blockIdx.x);
}

__global__ __launch_bounds__(kUpdNT) void chol_argmax_final_kernel(const double* __restrict__ pmax,
                                                                   const long long* __restrict__ pidx, int nparts,
                                                                   CholState* __restrict__ st) {
    __shared__ double sv[kUpdNT];
    __shared__ long long si[kUpdNT];
    double best = -1.0;
    long long bidx = 0x7fffffffffffffffLL;
    for (int e = threadIdx.x; e < nparts; e += kUpdNT) argmax_combine(best, bidx, pmax[e], pidx[e]);
    double v;
    long long i;
    block_argmax<kUpdNT>(best, bidx, sv, si, &v, &i);
    if (threadIdx.x == 0) {
        st->value = v;
        st->pivot = i;
    }
}

// ---------------------------------------------------------------------------
// Consumers of the device factors (no copy of the factors to the host)
// ---------------------------------------------------------------------------
// gamma[k] = sum_{s,l} conj(L[k][s*n + l]) D[l*n + s]: one workgroup per
// factor row, fixed-order reduction.
__global__ __launch_bounds__(256) void chol_gamma_kernel(const zc* const* __restrict__ chunks, int rows_per_chunk,
                                                         long long n, const zc* __restrict__ dens,
                                                         zc* __restrict__ gamma) {
    __shared__ zc red[256];
    const int k = blockIdx.x;
    const zc* row = factor_row(chunks, rows_per_chunk, n * n, k);
    zc acc{0.0, 0.0};
    for (long long x = threadIdx.x; x < n * n; x += 256) {
        const long long s = x / n, l = x - s * n;
        const zc v = row[x], d = dens[l * n + s];
        // conj(v) * d
        acc.x = fma(v.x, d.x, acc.x);
        acc.x = fma(v.y, d.y, acc.x);
        acc.y = fma(v.x, d.y, acc.y);
        acc.y = fma(-v.y, d.x, acc.y);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[threadIdx.x].x += red[threadIdx.x + w].x;
            red[threadIdx.x].y += red[threadIdx.x + w].y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) gamma[k] = red[0];
}

// out[m,a,l,s] (C order over the four ranges) = sum_k L[k][mu,nu] conj(L[k][sig,lam])
// with mu = r[0] + m, nu = r[2] + a, lam = r[4] + l, sig = r[6] + s.
__global__ __launch_bounds__(256) void chol_reconstruct_kernel(const zc* const* __restrict__ chunks,
                                                               int rows_per_chunk, long long n, int rank,
                                                               const long long* __restrict__ r,
                                                               zc* __restrict__ out) {
    const long long nm = r[1] - r[0], na = r[3] - r[2], nl = r[5] - r[4], ns = r[7] - r[6];
    const long long total = nm * na * nl * ns;
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= total) return;
    const long long s = t % ns, l = (t / ns) % nl, a = (t / (ns * nl)) % na, m = t / (ns * nl * na);
    const long long left = (r[0] + m) * n + (r[2] + a), right = (r[6] + s) * n + (r[4] + l);
    zc acc{0.0, 0.0};
    for (int k = 0; k < rank; ++k) {
        const zc* row = factor_row(chunks, rows_per_chunk, n * n, k);
        const zc u = row[left], v = row[right];
        // u * conj(v)
        acc.x = fma(u.x, v.x, acc.x);
        acc.x = fma(u.y, v.y, acc.x);
        acc.y = fma(u.y, v.x, acc.y);
        acc.y = fma(-u.x, v.y, acc.y);
    }
    out[t] = acc;
}

// ---------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------
int chol_partials(long long n2) { return (int)((n2 + kUpdNT * kUpdPerThread - 1) / (kUpdNT * kUpdPerThread)); }

hipError_t launch_chol_expand(const CholUnit* units, int nunits, const CholKet& ket, SeedQuartet* sq, WorkItem* wi,
                              hipStream_t s) {
    if (nunits == 0) return hipSuccess;
    chol_expand_kernel<<<dim3((unsigned)((nunits + 255) / 256)), dim3(256), 0, s>>>(units, nunits, ket, sq, wi);
    return hipGetLastError();
}

hipError_t launch_chol_topk(const double* res, long long n2, const double* pmax, const long long* pidx, int nb,
                            long long* cand, int* ncand, hipStream_t s) {
    if (nb < 1 || nb > kCholMaxBlock) return hipErrorInvalidValue;
    chol_topk_kernel<<<dim3(1), dim3(kTopNT), 0, s>>>(res, n2, pmax, pidx, chol_partials(n2), nb, cand, ncand);
    return hipGetLastError();
}

hipError_t launch_chol_coef(const zc* const* chunks, int rows_per_chunk, long long n2, int rank, const long long* piv,
                            int nnew, int nbt, zc* coef, hipStream_t s) {
    if (rank == 0) return hipSuccess;
    const long long total = (long long)rank * nbt;
    chol_coef_kernel<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>(chunks, rows_per_chunk, n2, rank, piv,
                                                                                nnew, nbt, coef);
    return hipGetLastError();
}

template <int NB, int NE, int U>
static hipError_t project(const zc* const* chunks, int rows_per_chunk, long long n2, int rank, const zc* coef, int nnew,
                          const int* slots, zc* proj, hipStream_t s) {
    const long long per = 256LL * NE;
    chol_project_kernel<NB, NE, U><<<dim3((unsigned)((n2 + per - 1) / per)), dim3(256), 0, s>>>(
        chunks, rows_per_chunk, n2, rank, coef, nnew, slots, proj);
    return hipGetLastError();
}

int chol_block_slots(int nb) {
    int t = 1;
    while (t < nb) t *= 2;
    return t;
}

hipError_t launch_chol_project(const zc* const* chunks, int rows_per_chunk, long long n2, int rank, const zc* coef,
                               int nnew, int nbt, const int* slots, zc* proj, hipStream_t s) {
    if (nnew == 0) return hipSuccess;
    switch (nbt) {
        case 1: return project<1, 4, 2>(chunks, rows_per_chunk, n2, rank, coef, nnew, slots, proj, s);
        case 2: return project<2, 4, 2>(chunks, rows_per_chunk, n2, rank, coef, nnew, slots, proj, s);
        case 4: return project<4, 2, 2>(chunks, rows_per_chunk, n2, rank, coef, nnew, slots, proj, s);
        case 8: return project<8, 2, 2>(chunks, rows_per_chunk, n2, rank, coef, nnew, slots, proj, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_chol_project_mfma(const zc* const* chunks, int rows_per_chunk, long long n2, int rows,
                                    const zc* coef, int nnew, int nbt, const int* slots, zc* proj, hipStream_t s) {
    if (nnew == 0) return hipSuccess;
    if (rows % 4 != 0) return hipErrorInvalidValue;
    constexpr int NJ = 4;
    const long long per = 4LL * 16 * NJ;  // entries per workgroup
    const dim3 grid((unsigned)((n2 + per - 1) / per));
    if (nbt == 16)
        chol_project_mfma_kernel<1, NJ><<<grid, dim3(256), 0, s>>>(chunks, rows_per_chunk, n2, rows, coef, nnew, slots,
                                                                    proj);
    else if (nbt == 32)
        chol_project_mfma_kernel<2, NJ><<<grid, dim3(256), 0, s>>>(chunks, rows_per_chunk, n2, rows, coef, nnew, slots,
                                                                    proj);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_chol_catchup_mfma(const zc* const* chunks, int rows_per_chunk, long long n2, int from, int to,
                                    long long pm, zc* proj, hipStream_t s) {
    if (from >= to) return hipSuccess;
    if (from % 4 != 0) return hipErrorInvalidValue;
    const long long per = 4LL * 16 * kCatchNJ;
    chol_catchup_mfma_kernel<<<dim3((unsigned)((n2 + per - 1) / per)), dim3(256), 0, s>>>(chunks, rows_per_chunk, n2,
                                                                                       from, to, pm, proj);
    return hipGetLastError();
}

hipError_t launch_chol_gamma(const zc* const* chunks, int rows_per_chunk, long long n, int rank, const zc* dens,
                             zc* gamma, hipStream_t s) {
    if (rank == 0) return hipSuccess;
    chol_gamma_kernel<<<dim3((unsigned)rank), dim3(256), 0, s>>>(chunks, rows_per_chunk, n, dens, gamma);
    return hipGetLastError();
}

hipError_t launch_chol_reconstruct(const zc* const* chunks, int rows_per_chunk, long long n, int rank,
                                   const long long* ranges, long long total, zc* out, hipStream_t s) {
    if (total == 0) return hipSuccess;
    chol_reconstruct_kernel<<<dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s>>>(chunks, rows_per_chunk, n,
                                                                                       rank, ranges, out);
    return hipGetLastError();
}

// Final residual of a run whose screened diagonal blocks were measured at
// the end (cholesky.py:283-288): out[e] = D[e] - sum_k |L_k[e]|^2 over the
// factor rows in order k = 0..rank-1, |L|^2 taken as hypot(re, im)^2, with
// no clamp and no zeroed pivots; pmax[block] = the block's largest value.
__global__ __launch_bounds__(256) void chol_remeasure_kernel(const zc* const* __restrict__ chunks,
                                                             int rows_per_chunk, long long n2, int rank,
                                                             const double* __restrict__ diag,
                                                             double* __restrict__ pmax) {
    __shared__ double red[256];
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    double v = -HUGE_VAL;
    if (e < n2) {
        double sum = 0.0;
        for (int k = 0; k < rank; ++k) {
            const zc x = factor_row(chunks, rows_per_chunk, n2, k)[e];
            const double a = hypot(x.x, x.y);
            sum += a * a;
        }
        v = diag[e] - sum;
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) pmax[blockIdx.x] = red[0];
}

hipError_t launch_chol_remeasure(const zc* const* chunks, int rows_per_chunk, long long n2, int rank,
                                 const double* diag, double* pmax, hipStream_t s) {
    if (n2 == 0) return hipSuccess;
    chol_remeasure_kernel<<<dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s>>>(chunks, rows_per_chunk, n2, rank,
                                                                                  diag, pmax);
    return hipGetLastError();
}

__global__ void chol_gather_kernel(const double* __restrict__ res, const long long* __restrict__ idx, int m,
                                   double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) out[i] = res[idx[i]];
}

hipError_t launch_chol_gather(const double* res, const long long* idx, int m, double* out, hipStream_t s) {
    if (m <= 0) return hipSuccess;
    chol_gather_kernel<<<dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s>>>(res, idx, m, out);
    return hipGetLastError();
}

hipError_t launch_chol_accept(const zc* column, const zc* pproj, int real, long long n, long long n2,
                              const CholState* st, double* res, zc* out_row, double psd_tol, int* flags, double* pmax,
                              long long* pidx, hipStream_t s) {
    chol_accept_kernel<<<dim3((unsigned)chol_partials(n2)), dim3(kUpdNT), 0, s>>>(
        column, pproj, real, n, n2, st, res, out_row, psd_tol, flags, pmax, pidx);
    return hipGetLastError();
}

hipError_t launch_chol_argmax(const double* res, long long n2, double* pmax, long long* pidx, CholState* st,
                              bool partial_first, hipStream_t s) {
    const int parts = chol_partials(n2);
    if (partial_first)
        chol_argmax_partial_kernel<<<dim3((unsigned)parts), dim3(kUpdNT), 0, s>>>(res, n2, pmax, pidx);
    chol_argmax_final_kernel<<<dim3(1), dim3(kUpdNT), 0, s>>>(pmax, pidx, parts, st);
    return hipGetLastError();
}

}  // namespace gh
