// Internal interface between the C ABI (gh_api.hip) and the kernel files.
#pragma once

#include "gh_common.h"

namespace gh {

struct ClassArgs {
    int l[4];
    int tile;  // ket components per pass
};

// Device-resident Cholesky (gh_cholesky.hip).
struct CholUnit {  // one bra shell pair (and contraction rows) of a pivot column
    int sh_a, sh_b;
    int xa, xb;
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
// Pivot blocks (gh_cholesky.hip): at most kCholMaxBlock candidates.
constexpr int kCholMaxBlock = 32;
// Candidate slots of a block of nb candidates (nb rounded up to a power of two).
int chol_block_slots(int nb);
hipError_t launch_chol_topk(const double* res, long long n2, const double* pmax, const long long* pidx, int nb,
                            long long* cand, int* ncand, hipStream_t s);
hipError_t launch_chol_coef(const zc* const* chunks, int rows_per_chunk, long long n2, int rank, const long long* piv,
                            int nnew, int nbt, zc* coef, hipStream_t s);
// Projections of the nnew new candidates (coefficient columns padded to
// nbt, a power of two <= kCholMaxBlock) into candidate slots slots[m].
hipError_t launch_chol_project(const zc* const* chunks, int rows_per_chunk, long long n2, int rank, const zc* coef,
                               int nnew, int nbt, const int* slots, zc* proj, hipStream_t s);
// Consumers of the device factors: gamma[k] = sum conj(L[k][s,l]) D[l,s]
// (one workgroup per factor row) and reconstruction of AO blocks.
hipError_t launch_chol_gamma(const zc* const* chunks, int rows_per_chunk, long long n, int rank, const zc* dens,
                             zc* gamma, hipStream_t s);
hipError_t launch_chol_reconstruct(const zc* const* chunks, int rows_per_chunk, long long n, int rank,
                                   const long long* ranges, long long total, zc* out, hipStream_t s);
// Matrix-core projections: rows (a multiple of 4) factor rows for nbt = 16 or
// 32 coefficient columns; the catch-up starts at a multiple of 4.
hipError_t launch_chol_project_mfma(const zc* const* chunks, int rows_per_chunk, long long n2, int rows,
                                    const zc* coef, int nnew, int nbt, const int* slots, zc* proj, hipStream_t s);
hipError_t launch_chol_catchup_mfma(const zc* const* chunks, int rows_per_chunk, long long n2, int from, int to,
                                    long long pm, zc* proj, hipStream_t s);
// out[i] = res[idx[i]] for i < m (residual entries the host pivot policy reads).
hipError_t launch_chol_gather(const double* res, const long long* idx, int m, double* out, hipStream_t s);
// pmax[(n2 + 255) / 256] = per-block max of D - sum_k |L_k|^2 (screened runs' final residual)
hipError_t launch_chol_remeasure(const zc* const* chunks, int rows_per_chunk, long long n2, int rank,
                                 const double* diag, double* pmax, hipStream_t s);
hipError_t launch_chol_accept(const zc* column, const zc* pproj, int real, long long n, long long n2,
                              const CholState* st, double* res, zc* out_row, double psd_tol, int* flags, double* pmax,
                              long long* pidx, hipStream_t s);
hipError_t launch_chol_argmax(const double* res, long long n2, double* pmax, long long* pidx, CholState* st,
                              bool partial_first, hipStream_t s);

// Nuclear-attraction matrix (nao x nao complex, every element written): one
// wavefront per ordered shell pair; nuclei as (Z, x, y, z) rows. Needs
// la + lb <= kNucMaxL for every pair.
constexpr int kNucMaxL = 8;
constexpr int kNucMaxComp = 4;  // component pairs per lane: (g, g) has 225
hipError_t launch_nuclear(int nshell, const DevShell* shells, const double* exps, const double* coefs,
                          const double* norms, const double* kappa, const double* nuclei, int nnuc,
                          const long long* ao_off, long long nao, zc* out, DevError* err, hipStream_t s);

// Class-specialized batched transform (gh_fast.h, gh_fast_*.hip, dispatch in gh_eri_fast.hip).
constexpr int kFastBraGroups = 4;  // waves running the bra step, each with a quarter of the bra components
constexpr int kFastMaxWaves = 8;   // largest transform workgroup
struct FastPairQuartet {  // one oriented shell quartet: pair-table and R0 bookkeeping
    int sh[4];
    long long unit_begin[2];  // first flat bra / ket primitive-pair unit
    long long table_off[2];   // zc offsets of its bra / ket product tables
    long long prim_begin;     // first flat primitive-quartet unit of the launch
    long long seed_off;       // zc offset of its seeds (L + 4 per primitive quartet)
    long long r0_off;         // zc offset of its R0 tables (nh(L) per primitive quartet)
    int kab, kcd;             // primitive pairs on the bra / ket side
};
struct FastItem {  // one oriented quartet x contraction rows
    int sh[4];
    int x[4];
    int conj, pad;
    long long r0_off;
    long long ebra_off, eket_off;
    long long out_base;
    long long stride[4];
};
struct FastLaunchArgs {
    const FastPairQuartet* pq;
    int npq;
    long long bra_units, ket_units, prim_units;
    const FastItem* items;
    int nitems;
    int waves;  // workgroup size of the transform in wavefronts
    const DevShell* shells;
    const double* exps;
    const double* coefs;
    const double* kappa;
    const zc* seeds;
    zc* r0;
    zc* etab;
    zc* out;
};
using FastLaunch = hipError_t (*)(const FastLaunchArgs& a, hipStream_t s);
struct FastClassInfo {
    bool available;
    int bra_ne, ket_ne;  // products per bra / ket primitive pair
    int nhab;            // bra Hermite indices (rows per bra primitive pair)
    int nhl;             // R0 entries per primitive quartet
    FastLaunch produce;  // pair tables and R0 tables (after the seeds)
    FastLaunch transform;  // the E.R.E transform (after produce)
};
FastClassInfo fast_class_info(int la, int lb, int lc, int ld);
// Transform workgroup size (wavefronts) for a bra side of nhab Hermite rows
// and kab primitive pairs; the kernel derives the same pass split from it.
// Waves per workgroup of the class-specialized transform for an item with kab
// bra primitive pairs: enough rows for every pair in one pass, at most `cap`
// waves (kFastBraGroups..kFastMaxWaves; 0 means kFastMaxWaves).
int fast_waves(int nhab, int kab, int cap = 0);

hipError_t upload_constants();
int transform_lds_bytes(const ClassArgs& ca);
// Transform entries per workgroup (W entries, output entries) for nt threads.
int transform_max_w(int nt);
int transform_max_o(int nt);

// Seeds of `total` primitive quartets. `scratch` holds seed_scratch_ints(total)
// ints (quadrature work lists); it is reused by the next launch in stream order.
constexpr int kBoysBins = 12;          // Boys phase-B lists: 11 quadrature bins, the power series
constexpr int kSeedScratchHead = 448;  // ints of counters before the phase-B lists
size_t seed_scratch_ints(long long total);
// E.R.E contraction on the FP64 matrix cores (gh_mfma.h) for bra / ket
// pair momenta (lab, lcd), lab >= lcd: whether an instance exists, the ket
// components one work item covers (items carry cd1 - cd0 <= that), and the
// launch (seeds as for launch_transform, one 512-thread workgroup per item)
// in compile-time variant var (GEMM unroll, LDS row padding; 0..3, the same
// integrals for every variant).
constexpr int kMfmaVariantCount = 4;
bool mfma_instance(int lab, int lcd);
int mfma_cd_tile(int lab);
hipError_t launch_mfma(int lab, int lcd, int var, const WorkItem* items, int n, const DevShell* shells, const double* exps,
                       const double* coefs, const double* norms, const double* kappa, const zc* seeds, zc* out,
                       hipStream_t s);
// Lane-per-primitive-quartet kernel of the smallest classes (gh_small.hip):
// la + lb + lc + ld <= kSmallMaxL, one 64-thread workgroup per item (a shell
// quartet x contraction rows, any ket range), seeds as for launch_transform.
constexpr int kSmallMaxL = 2;
bool small_class(int la, int lb, int lc, int ld);
hipError_t launch_small(int la, int lb, int lc, int ld, const WorkItem* items, int n, const DevShell* shells,
                        const double* exps, const double* coefs, const double* norms, const double* kappa,
                        const zc* seeds, zc* out, hipStream_t s);
// The split-FP32 form of the same contraction (FP32 hi/lo products on
// v_mfma_f32_16x16x4f32, compensated into FP64), instances (8, 8) and (6, 4).
bool mfma_split_instance(int lab, int lcd);
hipError_t launch_mfma_split(int lab, int lcd, const WorkItem* items, int n, const DevShell* shells,
                             const double* exps, const double* coefs, const double* norms, const double* kappa,
                             const zc* seeds, zc* out, hipStream_t s);
// p[0..n) = 0 by a kernel on stream s.
// ERI centre derivatives (gh_deriv.hip): one derivative quartet's block of
// 12 x size values formed from the shifted-momentum integrals in `t`.
struct DerivItem {
    long long out;       // output offset of the (4, 3, na, nb, nc, nd) block
    long long base;      // offset of (ab|cd) in t
    long long plus[4];   // offset of the slot-k (l + 1) block in t
    long long minus[4];  // offset of the slot-k (l - 1) block in t (unused for l = 0)
    long long size;      // na nb nc nd
    int n[4];            // AO counts of the four shells
    int l[4];
    int krow[4];         // contraction rows of the shells
    int prow[4];         // rows of the (l + 1) shells: krow, or 2 krow with the phase rows
    int first_tile, pad;
    double center[4][3];
};
hipError_t launch_deriv_assemble(const DerivItem* items, const int* tile_item, int ntiles, const zc* t, zc* out,
                                 const double B[3], bool phase, hipStream_t s);
hipError_t launch_zero_ints(int* p, int n, hipStream_t s);
hipError_t launch_seed(int nmax_bucket, const SeedQuartet* q, int nq, long long total, const DevShell* shells,
                       const double* exps, const double* kappa, zc* seeds, DevError* err, int* scratch,
                       hipStream_t s, const double* radial = nullptr);
hipError_t launch_transform(const ClassArgs& ca, const WorkItem* items, int nitems, const DevShell* shells,
                            const double* exps, const double* coefs, const double* norms, const double* kappa,
                            const zc* seeds, zc* out, hipStream_t s);
// Transform of items of several classes in one launch: every item carries
// its class and tile in WorkItem::cls; lds_bytes covers the largest; nt is the
// workgroup size the items were planned for (256, or 64: one wavefront). With
// `queue` (one device int, used only by launches on stream s) the launch is
// persistent: as many workgroups as fit on the device take items from a
// work queue in list order; without it, one workgroup per item.
hipError_t launch_transform_mixed(int nt, int lds_bytes, const WorkItem* items, int nitems, const DevShell* shells,
                                  const double* exps, const double* coefs, const double* norms, const double* kappa,
                                  const zc* seeds, zc* out, int* queue, hipStream_t s);
// Cholesky column ERIs for one (la, lb) group of bra shell pairs and the
// pivot's ket class (lc, ld) and component `comp`: one wavefront per pair.
// Needs la + lb + lc + ld <= kColumnMaxL; `items` as for launch_transform
// with a one-component ket range.
constexpr int kColumnMaxL = 16;
int chol_column_lds_bytes(int la, int lb, int lc, int ld, int comp);
hipError_t launch_chol_column(const ClassArgs& ca, int comp, const WorkItem* items, int nitems, const DevShell* shells,
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
