// Internal interface between the C ABI (gh_api.hip) and the kernel files.
#pragma once

#include "gh_common.h"

namespace gh {

struct ClassArgs {
    int l[4];
    int tile;  // ket components per pass
};

// Device-resident Cholesky (gh_cholesky.hip).
struct CholUnit {  // one bra shell pair of a pivot column
    int sh_a, sh_b;
    int lab, kab;
    long long p1;        // sum of kab over earlier units
    long long p2;        // sum of kab * lab over earlier units
    long long out_base;  // off_a * n + off_b
};
struct CholKet {  // the pivot's ket (sb, sa) and component
    int sh_c, sh_d;
    int lcd, kcd;
    int xc, xd;
    int comp, pad;
    long long n;
};
struct CholState {
    double value;      // current maximum residual
    long long pivot;   // its flat index (first maximum)
};

int chol_partials(long long n2);
hipError_t launch_chol_expand(const CholUnit* units, int nunits, const CholKet& ket, SeedQuartet* sq, WorkItem* wi,
                              hipStream_t s);
hipError_t launch_chol_gather(const zc* const* chunks, int rows_per_chunk, long long n2, int rank, const CholState* st,
                              zc* coef, hipStream_t s);
hipError_t launch_chol_update(const zc* column, const zc* const* chunks, int rows_per_chunk, long long n2, int rank,
                              const zc* coef, const CholState* st, double* res, zc* out_row, double psd_tol,
                              int* flags, double* pmax, long long* pidx, hipStream_t s);
hipError_t launch_chol_argmax(const double* res, long long n2, double* pmax, long long* pidx, CholState* st,
                              bool partial_first, hipStream_t s);

hipError_t upload_constants();
int transform_lds_bytes(const ClassArgs& ca);
int transform_max_w();
int transform_max_o();

hipError_t launch_seed(int nmax_bucket, const SeedQuartet* q, int nq, long long total, const DevShell* shells,
                       const double* exps, const double* kappa, zc* seeds, DevError* err, hipStream_t s);
hipError_t launch_transform(const ClassArgs& ca, const WorkItem* items, int nitems, const DevShell* shells,
                            const double* exps, const double* coefs, const double* norms, const double* kappa,
                            const zc* seeds, zc* out, hipStream_t s);
hipError_t launch_boys_probe(const zc* z, long long n, int nmax, int scaled, zc* out, int* region, double* errest,
                             int* segments, DevError* err, hipStream_t s);
hipError_t launch_tensor_scatter(const long long* block_off, const int* block_q, int nblocks, long long total,
                                 const DevShell* shells, const long long* ao_off, long long nao, const zc* blocks,
                                 zc* tensor, hipStream_t s);
hipError_t launch_diag_extract(const long long* block_off, const int* pairs, int npairs, const DevShell* shells,
                               const long long* ao_off, long long nao, const zc* blocks, double* diag,
                               hipStream_t s);
hipError_t launch_real_columns(const zc* cols, const int* owner, long long nao, long long npiv, double* out,
                               hipStream_t s);

}  // namespace gh
