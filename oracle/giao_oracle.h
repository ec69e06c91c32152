/*
 * CPU oracle for the London-orbital electron-repulsion path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of GIAOINT's
 * reference algorithm (per-Cartesian-component McMurchie-Davidson kernel with
 * the six-index complex R table, region-dispatched complex Boys function and
 * selected Cholesky integrals). Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as a checker or as the
 * timed CPU baseline; the product library never links it.
 *
 * Pinning status of each function (see oracle/README and the tests):
 *   or_boys ............ pinned: F_n(0)=1/(2n+1), recurrence residual,
 *                        conjugation, scaled/unscaled agreement, mpmath 1F1,
 *                        and the compiled reference (oracle/_ref).
 *   or_primitive_eri ... pinned: 192 frozen reference vectors
 *                        (tests/data/eri_recursive_m8.jsonl), ssss closed form,
 *                        pair-exchange / conjugate-reversal invariants.
 *   or_eri_batch ....... pinned: compiled reference, contracted = sum of
 *                        primitives, symmetry images.
 *   or_cholesky_* ...... pinned: compiled reference, dense Cholesky of W.
 *
 * Basis layout shared by every C entry point in this repository:
 *   nshell, centers[3*nshell] (bohr), L[nshell], nprim[nshell],
 *   ncontr[nshell], exps[sum nprim], coefs[sum ncontr*nprim] where the
 *   coefficients of shell s are row-major [contraction][primitive] and are
 *   already contraction-normalized (types.cpp:272-305 semantics).
 * Field layout: field[6] = {Bx, By, Bz, Ox, Oy, Oz} (atomic units, bohr).
 * Complex values are interleaved (re, im) doubles.
 */
#ifndef GIAO_ORACLE_H
#define GIAO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (mirroring the reference's C++ exception types). */
enum {
    OR_OK = 0,
    OR_INVALID_ARGUMENT = 1, /* std::invalid_argument */
    OR_LENGTH_ERROR = 2,     /* std::length_error */
    OR_OUT_OF_RANGE = 3,     /* std::out_of_range */
    OR_BOYS_ERROR = 4,       /* giao::BoysNumericalError */
    OR_RUNTIME_ERROR = 5
};

typedef struct {
    int nshell;
    const double* centers;
    const int* L;
    const int* nprim;
    const int* ncontr;
    const double* exps;
    const double* coefs;
} or_basis;

/* Last error message of the calling thread. */
const char* or_last_error(void);

/* types.cpp:207-219 */
double or_primitive_normalization(double exponent, int ax, int ay, int az);

/* types.cpp:272-305 (normalize=True path); coefs_out may alias coefs_in. */
int or_normalize_shell(int L, int nprim, int ncontr, const double* exps,
                       const double* coefs_in, double* coefs_out);

/* boys.cpp:245-294. out: 2*(nmax+1) doubles. region: 0 series, 1 adaptive
 * quadrature, 2 scaled quadrature, 3 positive asymptotic. */
int or_boys(double zr, double zi, int nmax, int scaled, double* out,
            int* region, double* error_estimate, int* segments);

/* eri.cpp:438-452 with EriKernel::compute (eri.cpp:67-384).
 * prim[i] = {exponent, cx, cy, cz, coefficient, normalized}, ang[i] = (x,y,z). */
int or_primitive_eri(const double prim[4][6], const int ang[4][3],
                     const double field[6], double out[2]);

/* Size in complex elements of the block of quartet (a,b,c,d). */
size_t or_quartet_size(const or_basis* basis, const uint32_t q[4]);

/* evaluate_eri_shell_quartets (eri.cpp:644-767) with a packing consumer:
 * block i is written at out + 2*offsets[i] (offsets in complex elements). */
int or_eri_batch(const or_basis* basis, const uint32_t* quartets, size_t n,
                 const double field[6], const uint64_t* offsets, double* out,
                 int threads);

/* compute_eri_tensor (eri.cpp:809-847): n^4 complex values. */
int or_eri_tensor(const or_basis* basis, const double field[6], double* out,
                  int threads);

/* canonicalize_shell_quartet (eri.cpp:414-429). Returns conjugate flag. */
int or_canonicalize(const uint32_t in[4], uint32_t out[4]);

/* EriSchwarzBounds (eri.cpp:587-620): out = nshell*nshell doubles. */
int or_schwarz_bounds(const or_basis* basis, const double field[6], double* out);

/* compute_eri_derivative_blocks (derivatives.cpp:1389-1415): offsets has
 * n + 1 entries (complex elements, offsets[0] = 0), block i is (4, 3, na,
 * nb, nc, nd) at out + 2*offsets[i]; out_len = offsets[n]. */
int or_eri_derivative_blocks(const or_basis* basis, const uint32_t* quartets, size_t n,
                             const double field[6], const uint64_t* offsets, double* out,
                             size_t out_len, int threads);

/* EriDerivativeSchwarzBounds values (derivatives.cpp:1307-1372): nshell*nshell. */
int or_eri_derivative_pair_bounds(const or_basis* basis, const double field[6], double* out);

/* cholesky.cpp:285-314. pairs: npairs*(a,b) with a<=b. out: n*n doubles. */
int or_cholesky_diagonal(const or_basis* basis, const uint64_t* pairs,
                         size_t npairs, const double field[6], double* out,
                         int threads);

/* nuclear.cpp:282-312 compute_nuclear_attraction_matrix: (nao, nao)
 * complex128, nuclei as charges[nnuc] and centers[nnuc][3]. */
int or_nuclear_matrix(const or_basis* basis, int nnuc, const double* charges, const double* centers,
                      const double field[6], double* out);

/* cholesky.cpp:203-282. pivots: npiv*(mu,nu). out: npiv*n*n values, complex
 * interleaved unless real_output (which requires exact zero field). */
int or_cholesky_columns(const or_basis* basis, const uint64_t* pivots,
                        size_t npiv, const double field[6], double* out,
                        int real_output, int threads);

#ifdef __cplusplus
}
#endif

#endif
