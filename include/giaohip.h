/*
 * giaohip: C ABI of the MI355X (gfx950) London-orbital ERI library.
 *
 * Plain pointers and sizes only. Every entry point returns a status code
 * (GIAOHIP_OK on success); the message of the last failure on a context is
 * read with giaohip_last_error. Status codes map to the exception types of the
 * reference C++ API (include/giao_integrals/eri.hpp, boys.hpp):
 *   GIAOHIP_INVALID_ARGUMENT -> std::invalid_argument / ValueError
 *   GIAOHIP_LENGTH_ERROR     -> std::length_error / ValueError
 *   GIAOHIP_OUT_OF_RANGE     -> std::out_of_range / IndexError
 *   GIAOHIP_BOYS_ERROR       -> giao::BoysNumericalError
 *   GIAOHIP_DEVICE_ERROR     -> HIP runtime failure (RuntimeError)
 *   GIAOHIP_UNSUPPORTED      -> a case the GPU kernels do not cover (RuntimeError)
 *
 * Basis layout (giaohip_basis_create): nshell shells; centers[3*nshell] in
 * bohr; L[s] total angular momentum; nprim[s]; ncontr[s]; exps = all
 * exponents shell by shell; coefs = per shell a row-major
 * [contraction][primitive] matrix, already contraction-normalized as
 * giao::Shell does (types.cpp:272-305). AOs of a shell are ordered
 * contraction-major, then x-major Cartesian component (types.hpp:78-80,
 * types.cpp:221-237); a basis orders its shells' AOs consecutively.
 * Field layout: field[6] = {Bx, By, Bz, Ox, Oy, Oz} (atomic units, bohr).
 * Complex values are interleaved (re, im) doubles (complex128 layout).
 *
 * A context owns one HIP stream on one device and serializes calls made on
 * it with a mutex; use one context per thread for concurrency.
 */
#ifndef GIAOHIP_H
#define GIAOHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GIAOHIP_OK = 0,
    GIAOHIP_INVALID_ARGUMENT = 1,
    GIAOHIP_LENGTH_ERROR = 2,
    GIAOHIP_OUT_OF_RANGE = 3,
    GIAOHIP_BOYS_ERROR = 4,
    GIAOHIP_DEVICE_ERROR = 5,
    GIAOHIP_UNSUPPORTED = 6,
    GIAOHIP_OUT_OF_MEMORY = 7, /* MemoryError: an output cap would be exceeded */
    GIAOHIP_RUNTIME_ERROR = 8  /* RuntimeError raised by the reference's driver */
};

/* ERI kernel variants (giaohip_set_option with GIAOHIP_OPTION_ERI_VARIANT). */
enum {
    GIAOHIP_ERI_PLAIN = 0,  /* generic FP64 transform kernel, one launch per class */
    GIAOHIP_ERI_BATCHED = 1 /* class-specialized FP64 kernels where available (default) */
};
enum {
    GIAOHIP_OPTION_ERI_VARIANT = 0,
    /* Work memory of one batched-path chunk (seeds, R0 tables, pair tables),
       in MiB, 16..65536; default 16384. Two chunks are in flight at once (one
       being produced, one transformed), so the path holds twice this much.
       Chunks of the generic transform hold up to twice this much of seeds,
       also two at once (four times in all). */
    GIAOHIP_OPTION_CHUNK_MIB = 1,
    /* Candidates per pivot block of giaohip_cholesky_run, 1..32; default 32.
       A block projects its candidates (the largest residuals) onto the
       earlier factors in one pass; consecutive pivots that are candidates
       need no further pass. Factors and pivots are bitwise identical for
       every value; it only trades passes over the factors against the
       projection work per pass. Read when a run starts. */
    GIAOHIP_OPTION_CHOL_BLOCK = 2,
    /* cholesky_eri's pivot_batch_size (python/giao_integrals/cholesky.py:
       193-206), >= 1; default 32. It only matters for complex runs on bases
       with general contractions, where the pivot may be taken from the
       latest candidate block instead of the global maximum. Read when a run
       starts. */
    GIAOHIP_OPTION_CHOL_PIVOT_BATCH = 3,
    /* Device staging of ERI outputs written to host memory, in MiB,
       1..65536; default 1024. Outputs are computed in groups of about this
       size; a group is copied to the host while the next one is computed,
       so two such buffers are held. */
    GIAOHIP_OPTION_STAGING_MIB = 4,
    /* Launch form of the generic transform kernel (classes without a
       specialized kernel, and every class under GIAOHIP_ERI_PLAIN), one of
       GIAOHIP_QUEUE_*; default GIAOHIP_QUEUE_NONE. Results are bitwise
       identical for every form. */
    GIAOHIP_OPTION_ERI_QUEUE = 5,
    /* Largest workgroup of the class-specialized transform, in waves:
       0 (meaning 8) or 4..8; default 4. A workgroup takes as many bra
       primitive pairs per pass as its rows hold, so the value sets the tile
       of bra pairs. */
    GIAOHIP_OPTION_FAST_WAVES = 6,
    /* Classes whose E.R.E contraction runs on the FP64 matrix cores
       (v_mfma_f64_16x16x4f64, dense complex Hermite products) under the
       batched variant, one of GIAOHIP_MFMA_*; default GIAOHIP_MFMA_DEFAULT.
       Instances exist for bra/ket pair momenta (LAB, LCD) with
       2 <= LCD <= LAB <= 6, and (7,7), (8,8), in either orientation. The
       plain variant never uses them. */
    GIAOHIP_OPTION_MFMA = 7,
    /* Scheduling of the batched path's producer kernels (seeds, pair
       tables, R0 tables), one of GIAOHIP_PIPELINE_*; default
       GIAOHIP_PIPELINE_OVERLAP. Results are bitwise identical for both. */
    GIAOHIP_OPTION_PIPELINE = 8,
    /* Largest total angular momentum la + lb + lc + ld of the classes whose
       generic-transform items run on one-wavefront workgroups (when their
       transform rows fit one wavefront); -1 puts every item on four-wavefront
       workgroups. Default 7. Results are bitwise identical for every value. */
    GIAOHIP_OPTION_WAVE_MAX_L = 9,
    /* Under GIAOHIP_MFMA_DEFAULT: smallest total angular momentum
       la + lb + lc + ld of the classes (without a class-specialized kernel,
       both pair momenta >= 2) whose contraction runs on the FP64 matrix
       cores; 0..33, default 9. */
    GIAOHIP_OPTION_MFMA_MIN_L = 10,
    /* Cholesky column launches of a pivot (expansion, Boys seeds unless
       buffered, the column kernels of every bra group over two streams) as
       hipGraphs: 1 (every pivot), 2 (pivots inside a candidate block only)
       or 0 (default: direct launches, measured faster over a whole run).
       One executable graph per launch topology, updated from each pivot's
       captured launches. Pivots and factors are bitwise identical for every
       value. */
    GIAOHIP_OPTION_CHOL_GRAPHS = 11,
    /* Compile-time variant of the FP64 matrix-core contraction kernels,
       0..3 (default 0): bit 0 flips the GEMM step unroll, bit 1 pads the
       LDS row stride of its W and ket-product tiles by 16 bytes. Results
       are bitwise identical for every variant (the autotuner picks one). */
    GIAOHIP_OPTION_MFMA_VARIANT = 12,
    /* Largest total angular momentum la + lb + lc + ld (-1..2, default 2)
       of the classes evaluated by the lane-per-primitive-quartet kernel (each
       lane one primitive quartet, the block in registers, one wavefront
       reduction) instead of the generic transform. */
    GIAOHIP_OPTION_SMALL_MAX_L = 13
};
enum {
    GIAOHIP_PIPELINE_OVERLAP = 0, /* on a second stream, beside the previous chunk's contraction */
    GIAOHIP_PIPELINE_SERIAL = 1   /* on the context stream, between the contractions */
};
enum {
    GIAOHIP_MFMA_OFF = 0,     /* FP64 VALU kernels only */
    GIAOHIP_MFMA_DEFAULT = 1, /* matrix cores where they are measured faster: total momentum >=
                                 GIAOHIP_OPTION_MFMA_MIN_L */
    GIAOHIP_MFMA_ALL = 2,     /* matrix cores for every class with an instance (precision studies) */
    GIAOHIP_MFMA_SPLIT32 = 3  /* as ALL, but (gg|gg) and (ff|dd)-type classes ((LAB, LCD) = (8, 8), (6, 4))
                                 in split FP32: hi/lo FP32 products on v_mfma_f32_16x16x4f32 accumulated in
                                 FP32 per step of 4 and compensated into FP64. A precision study: its errors
                                 (about 1e-7 relative) are far above the 1e-11 bar of the FP64 path. */
};
enum {
    GIAOHIP_QUEUE_NONE = 0,    /* one workgroup per item, one launch per LDS-size bucket */
    GIAOHIP_QUEUE_BUCKETS = 1, /* persistent device work queue per LDS-size bucket */
    GIAOHIP_QUEUE_SINGLE = 2   /* one persistent launch over every item (largest LDS of any) */
};

typedef struct giaohip_ctx giaohip_ctx;
typedef struct giaohip_basis giaohip_basis;
typedef struct giaohip_cholesky giaohip_cholesky;

/* Statistics of the last ERI call; times are device times from HIP events
 * recorded on the context's stream around the launches. Mirrors
 * giao::EriStatistics (eri.hpp:77-81) plus timing. */
typedef struct {
    uint64_t requested_quartets;
    uint64_t computed_quartets;
    uint64_t screened_quartets;
    uint64_t primitive_quartets;
    double seed_ms;      /* pair data + complex Boys + seeds kernel(s) */
    double transform_ms; /* R recurrence + E.R.E contraction kernel(s) */
    double total_ms;     /* everything on the stream, including copies */
    uint64_t launches;
    double host_ms;      /* wall time of the whole call, validation included */
    /* Split of transform_ms by kernel. Batched path: produce_ms = pair-table
     * and R0 kernels on the producer stream, contract_ms = E.R.E contraction
     * kernels on the context stream (giaohip_stream). Plain path: all of
     * transform_ms is contract_ms. contract_launches counts the contraction
     * launches those times cover. */
    double produce_ms;
    double contract_ms;
    uint64_t contract_launches;
} giaohip_stats;

/* --- context ---------------------------------------------------------- */
int giaohip_create(int device, giaohip_ctx** out);
int giaohip_destroy(giaohip_ctx* ctx);
/* Copies the last error message (NUL terminated) and returns its length. */
size_t giaohip_last_error(const giaohip_ctx* ctx, char* buffer, size_t size);
/* Stream (hipStream_t) the context launches on, for callers that time it. */
void* giaohip_stream(giaohip_ctx* ctx);
int giaohip_last_stats(const giaohip_ctx* ctx, giaohip_stats* out);
/* Context options; see GIAOHIP_OPTION_*. No counterpart in the reference. */
int giaohip_set_option(giaohip_ctx* ctx, int option, int value);
/* Device memory owned by the context, for resident outputs (bench, chaining). */
int giaohip_device_alloc(giaohip_ctx* ctx, size_t bytes, void** out);
int giaohip_device_free(giaohip_ctx* ctx, void* ptr);
int giaohip_memcpy_to_host(giaohip_ctx* ctx, void* dst, const void* src, size_t bytes);
int giaohip_synchronize(giaohip_ctx* ctx);

/* Host threads the library uses for its planning work (validation, work
 * lists, host copies): the hardware threads, at most 16. Replaces
 * openmp_max_threads (eri.cpp:863-869, decl. eri.hpp:145-146). */
int giaohip_host_threads(void);
/* Device allocations (hipMalloc calls) the library has made in this process
 * so far, on every device. A diagnostic with no counterpart in the
 * reference: it shows which calls run without allocating. */
unsigned long long giaohip_device_allocations(void);

/* --- basis (giao::Basis, types.hpp:108-124) ----------------------------- */
int giaohip_basis_create(giaohip_ctx* ctx, int nshell, const double* centers, const int* L, const int* nprim,
                         const int* ncontr, const double* exps, const double* coefs, giaohip_basis** out);
/* Replaces the shells of an existing basis (same layout as
 * giaohip_basis_create). Its device tables only grow: shells that fit the
 * tables of an earlier fill are uploaded without allocating device memory.
 * This is what keeps compute_eri's warm path free of device allocations,
 * as the reference's EriWorkspace keeps its own (eri.hpp:19-46,
 * eri.cpp:388-412). On failure the basis is left with no shells. */
int giaohip_basis_update(giaohip_ctx* ctx, giaohip_basis* basis, int nshell, const double* centers, const int* L,
                         const int* nprim, const int* ncontr, const double* exps, const double* coefs);
int giaohip_basis_destroy(giaohip_ctx* ctx, giaohip_basis* basis);
/* Number of AOs (Basis::ao_count, types.hpp:118). */
size_t giaohip_basis_ao_count(const giaohip_basis* basis);

/* --- host-side shell data (no GPU needed) ------------------------------
 * Contraction normalization of giao::Shell (types.cpp:239-306, normalize
 * policy) for n shells at once: exps/coefs_in/coefs_out follow the basis
 * layout above; coefs_out may alias coefs_in. */
int giaohip_normalize_shells(int n, const int* L, const int* nprim, const int* ncontr, const double* exps,
                             const double* coefs_in, double* coefs_out, char* message, size_t message_size);
/* primitive_normalization (types.cpp:207-219). */
double giaohip_primitive_normalization(double exponent, int ax, int ay, int az);

/* --- Boys function (replaces giao::compute_boys, boys.hpp:36-42) -------
 * Evaluates F_0..F_nmax for n arguments z (interleaved complex). out holds
 * n*(nmax+1) complex values; region (optional) holds giao::BoysRegion values
 * (0 series, 1 adaptive quadrature, 2 scaled quadrature, 3 asymptotic);
 * error_estimate and segments are optional per-argument diagnostics
 * (BoysDiagnostics, boys.hpp:24-29). Host pointers. */
int giaohip_boys(giaohip_ctx* ctx, const double* z, size_t n, int nmax, int scaled, double* out, int* region,
                 double* error_estimate, int* segments);

/* --- shell-quartet ERIs --------------------------------------------------
 * Replaces evaluate_eri_shell_quartets (eri.cpp:644-767, decl. eri.hpp:131-135)
 * with the packing consumer of _eri_batch (bindings.cpp:274-282,672-731):
 * quartets[4*i..] are shell indices; block i (C-order complex128 of shape
 * (ao_a, ao_b, ao_c, ao_d)) is written at out + 2*offsets[i] (offsets in
 * complex elements; pass NULL for the packed layout offsets[i] = sum of the
 * sizes of blocks 0..i-1). out is host memory unless out_on_device != 0, in
 * which case it is device memory from giaohip_device_alloc. */
int giaohip_eri_batch(giaohip_ctx* ctx, const giaohip_basis* basis, const uint32_t* quartets, size_t n,
                      const double field[6], const uint64_t* offsets, double* out, int out_on_device);

/* giaohip_eri_batch into caller-provided device memory without allocating:
 * no counterpart in the reference; it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same inputs, checks, errors and output
 * layout as giaohip_eri_batch with out_on_device = 1; every integral runs on
 * the plain FP64 transform kernel. Its work memory (seeds, descriptors,
 * error record) is a 64 MiB array compiled into the library, one per device,
 * resident until the process ends, and the last call's contents stay in it.
 * Calls that overlap on one device (several contexts on one GPU) race on it;
 * calls on one context are serialized. Limits: at most 8192 / 3 shells in the
 * basis, and each shell quartet's seeds (16 * (L + 4) bytes per primitive
 * quartet) within about 59 MiB; beyond them GIAOHIP_INVALID_ARGUMENT. Device
 * memory of the basis and of the output is the caller's (allocated earlier). */
int giaohip_eri_batch_noalloc(giaohip_ctx* ctx, const giaohip_basis* basis, const uint32_t* quartets, size_t n,
                              const double field[6], const uint64_t* offsets, double* out_device);

/* Replaces compute_eri_derivative_blocks (derivatives.cpp:1389-1415, decl.
 * derivatives.hpp:257-261; Python binding _eri_derivative_batch,
 * bindings.cpp:2255-2304): centre derivatives of the ERIs of n shell
 * quartets, London phase response included at fixed B and gauge origin.
 * Block i is (4, 3, na, nb, nc, nd) complex values (quartet slot, axis,
 * AOs), slots in quartet order even for repeated shells, at out +
 * 2*offsets[i]. offsets holds n + 1 entries (complex elements) with
 * offsets[0] = 0 and offsets[i+1] - offsets[i] = 12 * block size, or is NULL
 * (packed); out_len must equal offsets[n]. Errors, in the reference's order
 * and words: GIAOHIP_LENGTH_ERROR "derivative batch offsets have an incorrect
 * size or start" / "... do not match block sizes" / "derivative batch output
 * has an incorrect size", GIAOHIP_OUT_OF_RANGE "derivative shell quartet
 * index is out of range", then the ERI class errors of the shifted classes
 * (one more unit of angular momentum on each slot). out is device memory when
 * out_on_device, else host memory. */
int giaohip_eri_derivative_batch(giaohip_ctx* ctx, const giaohip_basis* basis, const uint32_t* quartets, size_t n,
                                 const double field[6], const uint64_t* offsets, double* out, uint64_t out_len,
                                 int out_on_device);

/* Replaces the values of EriDerivativeSchwarzBounds (derivatives.cpp:
 * 1307-1372, decl. derivatives.hpp:244-255): out = nshell*nshell doubles
 * (host); out[a*nshell+b] bounds every centre derivative of the pair density
 * (a b| by the triangle/Schwarz sum of the reference (shifted raw primitives,
 * London terms, |contraction coefficients|), padded by 128 eps and one ulp.
 * A quartet's bound is max(out[ab] Q[cd], Q[ab] out[cd]) with the ordinary
 * Schwarz factors Q of giaohip_schwarz_bounds (derivatives.cpp:1374-1383). */
int giaohip_eri_derivative_pair_bounds(giaohip_ctx* ctx, const giaohip_basis* basis, const double field[6],
                                       double* out);

/* Size in complex elements of the packed output of a quartet list. */
int giaohip_eri_batch_size(const giaohip_basis* basis, const uint32_t* quartets, size_t n, uint64_t* total);

/* Replaces EriSchwarzBounds (eri.cpp:587-620, decl. eri.hpp:83-101): out =
 * nshell*nshell doubles (host), out[a*nshell+b] = out[b*nshell+a] = the next
 * double above sqrt(max |(a_i b_j|b_j a_i)|) * (1 + 64 eps). */
int giaohip_schwarz_bounds(giaohip_ctx* ctx, const giaohip_basis* basis, const double field[6], double* out);

/* Screening step of evaluate_eri_shell_quartets (eri.cpp:652-688): checks the
 * threshold (finite, >= 0; > 0 needs bounds) and every shell index, then
 * copies the quartets with bounds[a,b]*bounds[c,d] >= threshold to kept (in
 * input order; kept may alias quartets) and their count to nkept. Host only. */
int giaohip_screen_quartets(giaohip_ctx* ctx, const giaohip_basis* basis, const uint32_t* quartets, size_t n,
                            const double* bounds, double threshold, uint32_t* kept, size_t* nkept);

/* Replaces compute_eri_tensor (eri.cpp:809-847, decl. eri.hpp:137-143):
 * the full (n,n,n,n) complex128 tensor from canonical quartets and their
 * four symmetry images. */
int giaohip_eri_tensor(giaohip_ctx* ctx, const giaohip_basis* basis, const double field[6], double* out,
                       int out_on_device);

/* Replaces primitive_eri (eri.cpp:438-452, decl. eri.hpp:114-121).
 * prim[i] = {exponent, cx, cy, cz, coefficient, normalized}; ang = powers. */
int giaohip_primitive_eri(giaohip_ctx* ctx, const double prim[4][6], const int ang[4][3], const double field[6],
                          double out[2]);

/* Replaces compute_nuclear_attraction_matrix (nuclear.cpp:282-312, decl.
 * nuclear.hpp:59-61; pybind11 `nuclear_attraction`, bindings.cpp:577-596):
 * V[mu,nu] = sum_C -Z_C <mu| 1/|r - C| |nu> with London orbitals, an
 * (nao, nao) complex128 matrix in host memory (row-major, every element
 * written). nuclei: charges[nnuc], centers[nnuc][3] in bohr. Shells through
 * l = 4 (GIAOHIP_UNSUPPORTED above; combined order above 32 is the
 * reference's invalid_argument). */
int giaohip_nuclear_attraction(giaohip_ctx* ctx, const giaohip_basis* basis, int nnuc, const double* charges,
                               const double* centers, const double field[6], double* out);

/* --- streamed Cholesky -------------------------------------------------
 * Replaces detail::cholesky_diagonal (cholesky.cpp:285-314, decl.
 * src/cholesky.hpp:9-11): pairs = npairs*(a,b) shell pairs with a <= b;
 * out = n*n doubles (host). */
int giaohip_cholesky_diagonal(giaohip_ctx* ctx, const giaohip_basis* basis, const uint64_t* pairs, size_t npairs,
                              const double field[6], double* out);

/* Replaces detail::cholesky_columns (cholesky.cpp:203-282,316-323, decl.
 * src/cholesky.hpp:12-17): pivots = npiv*(mu,nu) AO pairs sharing one shell
 * pair; out = npiv*n*n values, complex128 unless real_output (which requires
 * an exact zero field and gives float64). Host memory. */
int giaohip_cholesky_columns(giaohip_ctx* ctx, const giaohip_basis* basis, const uint64_t* pivots, size_t npiv,
                             const double field[6], double* out, int real_output);

/* Device-resident pivoted Cholesky of W[(mu,nu),(lam,sig)] = (nu mu|lam sig)
 * with strict global pivoting: the loop of cholesky_eri
 * (python/giao_integrals/cholesky.py:170-297) for the case where every
 * candidate set is the pivot itself (segmented shells, complex factors;
 * cholesky.py:193-247). Diagonal, argmax (first maximum in flat mu*n+nu
 * order), column ERIs, projection, scaling, residual update, PSD check and
 * clamp all run on the device; factors stay in device memory.
 * Requires ncontr == 1 for every shell (GIAOHIP_UNSUPPORTED otherwise).
 * As in cholesky.py:128-133,175-176,269-271, a basis on one centre is
 * factored at zero field (the London phase cancels), and zero-field or
 * one-centre runs use real arithmetic: each projected column is
 * symmetrized, (c[mu,nu] + c[nu,mu]) / 2, and the factors are real
 * (stored as complex128 with zero imaginary parts).
 * Stops when the largest residual diagonal is <= max(threshold,
 * 64*eps*max(diagonal)); max_rank caps the number of vectors
 * (GIAOHIP_OUT_OF_MEMORY beyond it, as MemoryError in the reference);
 * stop_after > 0 returns after that many pivots (timing of prefixes; the
 * result then reports converged = 0). */
int giaohip_cholesky_run(giaohip_ctx* ctx, const giaohip_basis* basis, const double field[6], double threshold,
                         uint64_t max_rank, uint64_t stop_after, giaohip_cholesky** out);
/* Continues a run stopped by stop_after from its device state until the
 * stopping test is met or the rank reaches stop_after (0: no limit). The
 * basis given to giaohip_cholesky_run must still exist. A run that ended with
 * an error cannot be continued. Replaces the reference driver's pivot loop
 * (python/giao_integrals/cholesky.py:183-297) continued in steps; it has no
 * counterpart call in the reference. */
int giaohip_cholesky_resume(giaohip_ctx* ctx, giaohip_cholesky* chol, uint64_t stop_after);
/* The final residual of a run on a basis with Schwarz-screened diagonal
 * blocks (cholesky_eri's screening_threshold, python/giao_integrals/
 * cholesky.py:138-154,283-288): every diagonal block is measured again and
 * max_residual = max(0, max over (mu,nu) of D[mu,nu] - sum_k |L_k[mu,nu]|^2),
 * unclamped entries, pivots not zeroed; also what giaohip_cholesky_info
 * reports afterwards. One pass over the factors. */
int giaohip_cholesky_remeasure(giaohip_ctx* ctx, giaohip_cholesky* chol, double* max_residual);
/* Rank, the measured largest residual diagonal, the stopping tolerance,
 * whether the stopping test was met, and device time split (ms). */
int giaohip_cholesky_info(const giaohip_cholesky* chol, uint64_t* rank, double* max_residual, double* tolerance,
                          int* converged, double* column_ms, double* update_ms, double* diagonal_ms);
/* Pivot-block statistics of a run: blocks started, factor rows read by their
 * projections (sum over the block starts that projected any candidate of the
 * rank rounded down to a multiple of 4, the projections' row groups; each
 * row is n*n complex), device time of candidate selection +
 * projection (ms, included in update_ms), the number of candidates
 * projected (candidates still open from the previous block keep their
 * projections), and candidate_rows = sum over blocks of candidates projected
 * x rank (each is n*n complex multiply-adds). Any output pointer may be NULL.
 * No counterpart in the reference (its projection is one GEMV per pivot). */
int giaohip_cholesky_blocks(const giaohip_cholesky* chol, uint64_t* blocks, uint64_t* factor_rows_read,
                            double* block_ms, uint64_t* projected, uint64_t* candidate_rows);
/* hipGraph use of a run's column launches (GIAOHIP_OPTION_CHOL_GRAPHS): graph
 * launches, graphs instantiated, in-place updates from a pivot's captured
 * launches. Any output pointer may be NULL. Replaces the per-pivot native
 * column batch calls of the reference driver
 * (python/giao_integrals/cholesky.py:249-254 calling
 * src/cholesky.cpp:203-282), which launch no graphs. */
int giaohip_cholesky_graphs(const giaohip_cholesky* chol, uint64_t* launches, uint64_t* instantiations,
                            uint64_t* updates);
/* Reconstructed ERIs from the device factors, (mu nu|lam sig) ~
 * sum_P L_P[mu,nu] conj(L_P[sig,lam]), on the AO ranges mu in
 * [ranges[0], ranges[1]), nu in [ranges[2], ranges[3]), lam in [ranges[4],
 * ranges[5]), sig in [ranges[6], ranges[7]); out is the C-order complex128
 * block (host). Replaces CholeskyERI.reconstruct for contiguous slices
 * (python/giao_integrals/cholesky.py:39-57) without copying the factors. */
int giaohip_cholesky_reconstruct(giaohip_ctx* ctx, const giaohip_cholesky* chol, const int64_t ranges[8],
                                 double* out);
/* Coulomb-type contraction with the device factors:
 * J[mu,nu] = sum_{lam,sig} (mu nu|lam sig) D[lam,sig] ~ sum_P L_P[mu,nu] gamma_P,
 * gamma_P = sum_{lam,sig} conj(L_P[sig,lam]) D[lam,sig]. density and out are
 * n*n complex128 (host). Two passes over the factors. No counterpart in the
 * reference (a device-resident consumer of its CholeskyERI factors). */
int giaohip_cholesky_coulomb(giaohip_ctx* ctx, const giaohip_cholesky* chol, const double* density, double* out);
/* Exchange-type contraction with the device factors:
 * K[mu,nu] = sum_{lam,sig} (mu lam|sig nu) D[lam,sig] ~ sum_P (L_P D L_P^H)[mu,nu]
 * (the factor convention of giaohip_cholesky_reconstruct); two complex
 * GEMMs of order n per factor (rocBLAS zgemm on the context's stream).
 * density and out are n*n complex128 (host). No counterpart in the reference
 * (a device-resident consumer of its CholeskyERI factors). */
int giaohip_cholesky_exchange(giaohip_ctx* ctx, const giaohip_cholesky* chol, const double* density, double* out);
/* Pivot AO pairs (mu, nu), rank*2 int64 values. */
int giaohip_cholesky_pivots(const giaohip_cholesky* chol, int64_t* out);
/* Factor rows first..first+count-1 as complex128 (count*n*n values, host). */
int giaohip_cholesky_vectors(giaohip_ctx* ctx, const giaohip_cholesky* chol, uint64_t first, uint64_t count,
                             double* out);
int giaohip_cholesky_destroy(giaohip_ctx* ctx, giaohip_cholesky* chol);

#ifdef __cplusplus
}
#endif

#endif
