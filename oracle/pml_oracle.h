/*
 * CPU oracle for polynomial-matrix arithmetic over Z/pZ (word-size p).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked or
 * loaded by the test-suite, by the smoke check and by the CPU-baseline leg of
 * the benchmark, never by the product library.  It restates, from the
 * mathematics and from the reference's algorithms, what PML computes:
 *
 *   - polynomial matrix product C = A*B (flint-extras/src/nmod_poly_mat_extra/
 *     mul.c:19-97 and its evaluation back-ends); the product is unique, so any
 *     exact algorithm is bit-exact with the reference;
 *   - middle product (A*B mod x^nhi) div x^nlo (mulmid.c:499-559, the test
 *     definition in test/t-mulmid.c:26-34);
 *   - left nullspace in compact form (nmod_mat_extra/left_nullspace.c:42-102),
 *     derived from the unique reduced row echelon form of A^T;
 *   - M-Basis (nmod_mat_poly_extra/nmod_mat_poly_mbasis.c:116-378) and
 *     PM-Basis (nmod_poly_mat_extra/approximant_basis.c:33-69);
 *   - the approximant-basis certificate (nmod_poly_mat_extra/verification.c:
 *     21-116).
 *
 * Layout conventions (all arrays row-major, entry-major):
 *   a polynomial matrix with r rows, c columns and padded length len is a
 *   uint64_t array of r*c*len words, entry (i,j) coefficient d at
 *   a[(i*c + j)*len + d]; coefficients are in [0,p).  Lengths are "padded":
 *   trailing zero coefficients are allowed, and the normalised length of an
 *   entry is the index of its last nonzero coefficient plus one.
 *
 * Pinning (how each function is checked without the reference, which needs
 * FLINT/NTL and cannot be built here): see the header of each group below.
 */
#ifndef PML_ORACLE_H
#define PML_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- scalar arithmetic (exact, via 128-bit products) ---------------- */
uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p);
uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p);
uint64_t orc_invmod(uint64_t a, uint64_t p);            /* p prime, a != 0 */
int orc_is_prime(uint64_t n);                           /* deterministic for n < 2^64 */
int orc_two_adic_valuation(uint64_t n);                 /* v_2(n), n > 0 */
/* Primitive 2^k-th root of unity: g^((p-1)/2^k) for the smallest g >= 2 that
 * yields order exactly 2^k. Returns 0 if 2^k does not divide p-1. */
uint64_t orc_root_of_unity_2k(uint64_t p, int k);

/* ---------------- threads ---------------- */
/* Host threads of the product, evaluation and transform loops (OpenMP);
 * n <= 0 restores OpenMP's default (OMP_NUM_THREADS, else all cores). */
void orc_set_threads(int n);
int orc_threads(void);

/* ---------------- seeded generator ---------------- */
/* SplitMix64 stream; uniform residues in [0,p) by rejection on the top
 * bitlen(p) bits. */
uint64_t orc_splitmix64(uint64_t *state);
void orc_rand_uniform(uint64_t *out, int64_t n, uint64_t p, uint64_t seed);

/* ---------------- polynomial-matrix product ---------------- */
/* Schoolbook (closed form): C[i][j] = sum_k A[i][k]*B[k][j]; C has padded
 * length lenA+lenB-1 (0 if either length is 0). */
void orc_polmat_mul_school(uint64_t *C, const uint64_t *A, const uint64_t *B,
                           int64_t m, int64_t k, int64_t n,
                           int64_t lenA, int64_t lenB, uint64_t p);
/* Same result through a radix-2 NTT (needs 2^ceil(log2(lenA+lenB-1)) | p-1
 * and p < 2^63): decimation in frequency forward, decimation in time inverse,
 * exact 128-bit pointwise sums; returns 0 on success, -1 if the prime does not
 * allow it. */
int orc_polmat_mul_ntt(uint64_t *C, const uint64_t *A, const uint64_t *B,
                       int64_t m, int64_t k, int64_t n,
                       int64_t lenA, int64_t lenB, uint64_t p);
/* Dispatches: NTT when possible, schoolbook otherwise. */
void orc_polmat_mul(uint64_t *C, const uint64_t *A, const uint64_t *B,
                    int64_t m, int64_t k, int64_t n,
                    int64_t lenA, int64_t lenB, uint64_t p);
/* Middle product: out has padded length max(nhi-nlo,0); out = (A*B mod x^nhi) div x^nlo. */
void orc_polmat_mulmid(uint64_t *out, const uint64_t *A, const uint64_t *B,
                       int64_t m, int64_t k, int64_t n,
                       int64_t lenA, int64_t lenB, int64_t nlo, int64_t nhi, uint64_t p);
/* Evaluate every entry at x (Horner) -> r*c matrix. */
void orc_polmat_eval(uint64_t *out, const uint64_t *A, int64_t r, int64_t c, int64_t len,
                     uint64_t x, uint64_t p);
/* Max normalised length over all entries. */
int64_t orc_polmat_max_length(const uint64_t *A, int64_t r, int64_t c, int64_t len);

/* ---------------- constant matrices ---------------- */
/* Reduced row echelon form in place (unique); returns rank, pivot columns in pivcols. */
int64_t orc_rref(uint64_t *M, int64_t r, int64_t c, uint64_t p, int64_t *pivcols);
/* Left nullspace in compact form, convention of left_nullspace.c:42-102:
 * A is m x n.  Returns nullity.  X receives nullity x rank (row-major),
 * permutation receives m entries (row rank profile, then complement). */
int64_t orc_left_nullspace_compact(uint64_t *X, int64_t *permutation,
                                   const uint64_t *A, int64_t m, int64_t n, uint64_t p);
int64_t orc_rank(const uint64_t *A, int64_t r, int64_t c, uint64_t p);

/* ---------------- approximant bases ---------------- */
/* M-Basis (iterative) and PM-Basis (divide and conquer).  F is m x n with
 * padded length lenF.  On return *P is a malloc'ed m x m polynomial matrix of
 * padded length *lenP (caller frees with orc_free) and shift[0..m) holds the
 * shift-shifted row degree of P. */
void orc_mbasis(uint64_t **P, int64_t *lenP, int64_t *shift,
                const uint64_t *F, int64_t m, int64_t n, int64_t lenF,
                int64_t order, uint64_t p);
void orc_pmbasis(uint64_t **P, int64_t *lenP, int64_t *shift,
                 const uint64_t *F, int64_t m, int64_t n, int64_t lenF,
                 int64_t order, uint64_t p);
/* Certificate of verification.c:21-116 (orientation ROW_LOWER).  Returns a
 * bit mask of failed checks (0 = approximant basis):
 *   1 = not ordered weak Popov, 2 = residual not 0 mod x^order,
 *   4 = deg det(P) differs from the sum D of the diagonal degrees (exact, as
 *       verification.c:73-84 tests it; undecided, so not set, for a P that is
 *       not weak Popov when p <= sum of row degrees or the work is large),
 *   8 = [P(0) | C] not full rank.  `seed` is unused (kept for callers). */
int orc_is_approximant_basis(const uint64_t *P, int64_t lenP,
                             const uint64_t *F, int64_t m, int64_t n, int64_t lenF,
                             int64_t order, const int64_t *shift, uint64_t p,
                             uint64_t seed);

void orc_free(void *ptr);

#ifdef __cplusplus
}
#endif

#endif
