// gfx950 kernels of the device-resident pivoted Cholesky of the ordered-pair
// Coulomb matrix W[(mu,nu),(lam,sig)] = (nu mu|lam sig)
// (python/giao_integrals/cholesky.py:60-297, strict global pivoting).
//
// Per pivot p = (mu_p, nu_p):
//   expand_kernel   writes the seed quartets and transform work items of the
//                   raw column (mu' nu'|nu_p mu_p) for every bra shell pair from
//                   a static bra table and the pivot's ket (cholesky.cpp:203-282),
//                   so no per-pivot host planning is needed;
//   seed/transform  the ERI kernels of gh_eri.hip (single ket component);
//   gather_kernel   c_k = conj(L[k][p]) for the k < rank existing factors;
//   update_kernel   column -= sum_k c_k L[k] (memory-bound GEMV over the
//                   factor rows), L[rank] = column / sqrt(res[p]),
//                   res -= |L[rank]|^2 with the PSD check, clamp at 0,
//                   res[p] = 0, and per-block (max, first index) partials;
//   argmax kernels  first maximum in flat index order (np.argmax semantics).
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
    w.x[0] = 0;
    w.x[1] = 0;
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
constexpr int kCoefChunk = 1024;

__device__ inline const zc* factor_row(const zc* const* chunks, int rows_per_chunk, long long n2, int k) {
    return chunks[k / rows_per_chunk] + (long long)(k % rows_per_chunk) * n2;
}

__global__ void chol_gather_kernel(const zc* const* __restrict__ chunks, int rows_per_chunk, long long n2, int rank,
                                   const CholState* __restrict__ st, zc* __restrict__ coef) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= rank) return;
    const zc v = factor_row(chunks, rows_per_chunk, n2, k)[st->pivot];
    coef[k] = zc{v.x, -v.y};
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

__global__ __launch_bounds__(kUpdNT) void chol_update_kernel(const zc* __restrict__ column,
                                                             const zc* const* __restrict__ chunks, int rows_per_chunk,
                                                             long long n2, int rank, const zc* __restrict__ coef,
                                                             const CholState* __restrict__ st,
                                                             double* __restrict__ res, zc* __restrict__ out_row,
                                                             double psd_tol, int* __restrict__ flags,
                                                             double* __restrict__ pmax, long long* __restrict__ pidx) {
    __shared__ zc sc[kCoefChunk];
    __shared__ double sv[kUpdNT];
    __shared__ long long si[kUpdNT];
    const int tid = threadIdx.x;
    const long long base = (long long)blockIdx.x * kUpdNT * kUpdPerThread;
    zc acc[kUpdPerThread];
    long long j[kUpdPerThread];
#pragma unroll
    for (int e = 0; e < kUpdPerThread; ++e) {
        j[e] = base + e * kUpdNT + tid;
        acc[e] = j[e] < n2 ? column[j[e]] : zc{0.0, 0.0};
    }
    for (int k0 = 0; k0 < rank; k0 += kCoefChunk) {
        const int cnt = min(kCoefChunk, rank - k0);
        __syncthreads();
        for (int k = tid; k < cnt; k += kUpdNT) sc[k] = coef[k0 + k];
        __syncthreads();
        int k = 0;
        for (; k + 4 <= cnt; k += 4) {
            zc v[4][kUpdPerThread];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const zc* row = factor_row(chunks, rows_per_chunk, n2, k0 + k + kk);
#pragma unroll
                for (int e = 0; e < kUpdPerThread; ++e) v[kk][e] = j[e] < n2 ? row[j[e]] : zc{0.0, 0.0};
            }
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const zc c = sc[k + kk];
#pragma unroll
                for (int e = 0; e < kUpdPerThread; ++e) {
                    acc[e].x -= c.x * v[kk][e].x - c.y * v[kk][e].y;
                    acc[e].y -= c.x * v[kk][e].y + c.y * v[kk][e].x;
                }
            }
        }
        for (; k < cnt; ++k) {
            const zc* row = factor_row(chunks, rows_per_chunk, n2, k0 + k);
            const zc c = sc[k];
#pragma unroll
            for (int e = 0; e < kUpdPerThread; ++e) {
                if (j[e] < n2) {
                    const zc v = row[j[e]];
                    acc[e].x -= c.x * v.x - c.y * v.y;
                    acc[e].y -= c.x * v.y + c.y * v.x;
                }
            }
        }
    }
    const double root = sqrt(st->value);
    const long long p = st->pivot;
    double best = -1.0;
    long long bidx = 0x7fffffffffffffffLL;
#pragma unroll
    for (int e = 0; e < kUpdPerThread; ++e) {
        if (j[e] < n2) {
            const zc val{acc[e].x / root, acc[e].y / root};
            out_row[j[e]] = val;
            double nr = res[j[e]] - (val.x * val.x + val.y * val.y);
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
// Launch wrappers
// ---------------------------------------------------------------------------
int chol_partials(long long n2) { return (int)((n2 + kUpdNT * kUpdPerThread - 1) / (kUpdNT * kUpdPerThread)); }

hipError_t launch_chol_expand(const CholUnit* units, int nunits, const CholKet& ket, SeedQuartet* sq, WorkItem* wi,
                              hipStream_t s) {
    if (nunits == 0) return hipSuccess;
    chol_expand_kernel<<<dim3((unsigned)((nunits + 255) / 256)), dim3(256), 0, s>>>(units, nunits, ket, sq, wi);
    return hipGetLastError();
}

hipError_t launch_chol_gather(const zc* const* chunks, int rows_per_chunk, long long n2, int rank, const CholState* st,
                              zc* coef, hipStream_t s) {
    if (rank == 0) return hipSuccess;
    chol_gather_kernel<<<dim3((unsigned)((rank + 255) / 256)), dim3(256), 0, s>>>(chunks, rows_per_chunk, n2, rank, st,
                                                                                 coef);
    return hipGetLastError();
}

hipError_t launch_chol_update(const zc* column, const zc* const* chunks, int rows_per_chunk, long long n2, int rank,
                              const zc* coef, const CholState* st, double* res, zc* out_row, double psd_tol,
                              int* flags, double* pmax, long long* pidx, hipStream_t s) {
    chol_update_kernel<<<dim3((unsigned)chol_partials(n2)), dim3(kUpdNT), 0, s>>>(
        column, chunks, rows_per_chunk, n2, rank, coef, st, res, out_row, psd_tol, flags, pmax, pidx);
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
