/*
 * pmlgpu_flint.h: FLINT-shaped drop-in for PML's polynomial-matrix
 * multiplication, middle product and approximant bases, computed on the GPU
 * by libpmlgpu.so.
 *
 * Reference interfaces replaced (vneiger/pml, flint-extras/src):
 *   nmod_poly_mat_multiply          nmod_poly_mat_multiply.h:92   (mul.c:19-97)
 *   nmod_poly_mat_mul_geometric     nmod_poly_mat_multiply.h:80   (mul_geometric.c:121-157)
 *   nmod_poly_mat_mul_vandermonde1  nmod_poly_mat_multiply.h:63   (mul_vandermonde.c:71-209)
 *   nmod_poly_mat_mul_vandermonde2  nmod_poly_mat_multiply.h:64   (mul_vandermonde.c:268-503)
 *   nmod_poly_mat_mul_waksman       nmod_poly_mat_multiply.h:72   (mul_waksman.c:17-157)
 *   nmod_poly_mat_mulmid            nmod_poly_mat_multiply.h:122-124 (mulmid.c:499-559)
 *   _nmod_poly_mat_mulmid           nmod_poly_mat_multiply.h:128-131 (mulmid.c:470-497)
 *   nmod_poly_mat_pmbasis           nmod_poly_mat_approximant.h:282-285 (approximant_basis.c:33-69)
 *   nmod_poly_mat_mbasis            nmod_poly_mat_approximant.h:321-324 (approximant_basis.c:18-31)
 *   nmod_poly_mat_is_approximant_basis nmod_poly_mat_approximant.h:194-198 (verification.c:21-116)
 * The product is unique, so the evaluation back-ends (geometric, Vandermonde,
 * Waksman) all return what nmod_poly_mat_multiply returns; they are kept as
 * names only.
 *
 * Conventions kept from the reference: `res` is initialised by the caller
 * with the right dimensions and modulus; entries are grown and normalised by
 * the callee; the output may alias an input; zero-length operands give zero;
 * `shift` is in/out for the bases.  Errors (a modulus or length the device
 * path does not support, a device failure) abort with a message, as
 * flint_throw does; there is no CPU fallback.
 *
 * Types: with PMLG_USE_FLINT defined, FLINT's own nmod_poly_mat.h is used
 * (source compatibility; the struct layout is FLINT's).  Otherwise a subset
 * with FLINT's field names is declared here, with the layout of FLINT >= 3.3
 * (nmod_poly_mat_struct = entries, r, c, modulus; nmod_poly_struct = coeffs,
 * alloc, length, mod), so that matrices built by a FLINT 3.4 caller are read
 * as they are; the few FLINT memory-management calls the drop-in needs are
 * provided as pmlg_-prefixed functions mapped to FLINT's names by macros.
 */
#ifndef PMLGPU_FLINT_H
#define PMLGPU_FLINT_H

#include "pmlgpu.h"

#ifdef PMLG_USE_FLINT
#include <flint/nmod_poly.h>
#include <flint/nmod_poly_mat.h>
#else

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned long ulong;
typedef long slong;

typedef struct {
    ulong n;
    ulong ninv;
    ulong norm;
} nmod_t;

typedef struct {
    ulong *coeffs;
    slong alloc;
    slong length;
    nmod_t mod;
} nmod_poly_struct;
typedef nmod_poly_struct nmod_poly_t[1];

/* FLINT >= 3.3 layout (the reference needs FLINT >= 3.4, README.md:9):
 * entries row-major and contiguous, no `rows` array; entry (i, j) is
 * entries + i*c + j (nmod_poly_mat_utils.c:34-49 and
 * nmod_mat_poly_mbasis.c:321-345 take this branch from FLINT 3.3 on).
 * 32 bytes: entries at 0, r at 8, c at 16, modulus at 24. */
typedef struct {
    nmod_poly_struct *entries;
    slong r;
    slong c;
    ulong modulus;
} nmod_poly_mat_struct;
typedef nmod_poly_mat_struct nmod_poly_mat_t[1];

#define nmod_poly_mat_entry(mat, i, j) ((mat)->entries + (i) * (mat)->c + (j))
#define nmod_poly_mat_nrows(mat) ((mat)->r)
#define nmod_poly_mat_ncols(mat) ((mat)->c)

void pmlg_nmod_poly_init(nmod_poly_t poly, ulong n);
void pmlg_nmod_poly_clear(nmod_poly_t poly);
void pmlg_nmod_poly_fit_length(nmod_poly_t poly, slong len);
void pmlg_nmod_poly_set_coeff_ui(nmod_poly_t poly, slong j, ulong c);
ulong pmlg_nmod_poly_get_coeff_ui(const nmod_poly_t poly, slong j);
void pmlg_nmod_poly_mat_init(nmod_poly_mat_t mat, slong rows, slong cols, ulong n);
void pmlg_nmod_poly_mat_clear(nmod_poly_mat_t mat);
void pmlg_nmod_poly_mat_zero(nmod_poly_mat_t mat);
slong pmlg_nmod_poly_mat_max_length(const nmod_poly_mat_t mat);
int pmlg_nmod_poly_mat_equal(const nmod_poly_mat_t a, const nmod_poly_mat_t b);

#define nmod_poly_init pmlg_nmod_poly_init
#define nmod_poly_clear pmlg_nmod_poly_clear
#define nmod_poly_fit_length pmlg_nmod_poly_fit_length
#define nmod_poly_set_coeff_ui pmlg_nmod_poly_set_coeff_ui
#define nmod_poly_get_coeff_ui pmlg_nmod_poly_get_coeff_ui
#define nmod_poly_mat_init pmlg_nmod_poly_mat_init
#define nmod_poly_mat_clear pmlg_nmod_poly_mat_clear
#define nmod_poly_mat_zero pmlg_nmod_poly_mat_zero
#define nmod_poly_mat_max_length pmlg_nmod_poly_mat_max_length
#define nmod_poly_mat_equal pmlg_nmod_poly_mat_equal

#ifdef __cplusplus
}
#endif

#endif /* PMLG_USE_FLINT */

#ifdef __cplusplus
extern "C" {
#endif

/* orientation_t of nmod_poly_mat_forms.h:100-103 */
typedef enum { COL_UPPER = 0, COL_LOWER = 1, ROW_UPPER = 2, ROW_LOWER = 3 } orientation_t;
/* poly_mat_form_t of nmod_poly_mat_forms.h:57-66 */
typedef enum {
    NONE = 0,
    REDUCED = 1,
    WEAK_POPOV = 2,
    ORD_WEAK_POPOV = 3,
    POPOV = 4,
    ECHELON = 5,
    HERMITE = 6
} poly_mat_form_t;

void nmod_poly_mat_multiply(nmod_poly_mat_t res, const nmod_poly_mat_t A, const nmod_poly_mat_t B);
void nmod_poly_mat_mul_geometric(nmod_poly_mat_t res, const nmod_poly_mat_t A,
                                 const nmod_poly_mat_t B);
void nmod_poly_mat_mul_vandermonde1(nmod_poly_mat_t res, const nmod_poly_mat_t A,
                                    const nmod_poly_mat_t B);
void nmod_poly_mat_mul_vandermonde2(nmod_poly_mat_t res, const nmod_poly_mat_t A,
                                    const nmod_poly_mat_t B);
void nmod_poly_mat_mul_waksman(nmod_poly_mat_t res, const nmod_poly_mat_t A,
                               const nmod_poly_mat_t B);
void nmod_poly_mat_mulmid(nmod_poly_mat_t res, const nmod_poly_mat_t A, const nmod_poly_mat_t B,
                          slong nlo, slong nhi);
void _nmod_poly_mat_mulmid(nmod_poly_mat_t res, const nmod_poly_mat_t A, slong len1,
                           const nmod_poly_mat_t B, slong len2, slong nlo, slong nhi);
void nmod_poly_mat_pmbasis(nmod_poly_mat_t appbas, slong *shift, const nmod_poly_mat_t F,
                           slong order);
void nmod_poly_mat_mbasis(nmod_poly_mat_t appbas, slong *shift, const nmod_poly_mat_t F,
                          ulong order);
/* Certificate (ROW_LOWER only, as the reference): square; shift-ordered weak
 * Popov; appbas*F = 0 mod x^order (residual on the device); deg det(appbas)
 * equal to the sum D of the diagonal degrees, as the reference tests it
 * (verification.c:73-84: the degree only, not that det is a monomial); and
 * [appbas(0) | coefficient `order` of appbas*F] of full rank (rank on the
 * device).  For a square shift-ordered weak Popov matrix deg det = D holds by
 * construction (its shifted leading matrix is triangular with a nonzero
 * diagonal); otherwise the degree is found exactly by device determinants at
 * sum-of-row-degrees + 1 points and interpolation, when p exceeds that count
 * (else its message is not printed).  The result is the reference's. */
int nmod_poly_mat_is_approximant_basis(const nmod_poly_mat_t appbas, const nmod_poly_mat_t F,
                                       slong order, const slong *shift, orientation_t orient);

/* Number of device contexts the drop-in keeps (one per modulus, at most 32,
 * least recently used replaced first; no reference counterpart, a
 * diagnostic for tests). */
int pmlg_dropin_cached_contexts(void);

/* Minimal kernel bases (nmod_poly_mat_kernel.h:81-105), computed on the
 * device (pmlg_kernel_basis): the basis occupies the first `nullity` rows
 * (left kernels) or columns (right kernels) of ker, which keeps its shape;
 * pivind and shift (in/out, may be NULL for nmod_poly_mat_kernel) hold the
 * pivot indices and shifted degrees of those rows/columns.  Only the shifted
 * ordered weak Popov form is produced (form <= ORD_WEAK_POPOV), as in the
 * reference. */
slong nmod_poly_mat_kernel(nmod_poly_mat_t ker, slong *pivind, slong *shift,
                           const nmod_poly_mat_t pmat, poly_mat_form_t form, orientation_t orient);
slong nmod_poly_mat_kernel_via_approx(nmod_poly_mat_t ker, slong *pivind, slong *shift,
                                      const nmod_poly_mat_t pmat);
slong nmod_poly_mat_kernel_zls_approx(nmod_poly_mat_t ker, slong *pivind, slong *shift,
                                      nmod_poly_mat_t pmat);
/* verification.c:119-215 (nmod_poly_mat_kernel.h:144-148): dimensions,
 * nullity = dimension - rank (exact rank by device evaluations), the form
 * (NONE, WEAK_POPOV or ORD_WEAK_POPOV) and ker*mat == 0 or mat*ker == 0 on the
 * device.  Generation is not checked, as in the reference. */
int nmod_poly_mat_is_kernel(const nmod_poly_mat_t ker, const slong *shift,
                            const nmod_poly_mat_t mat, poly_mat_form_t form, orientation_t orient);

#ifdef __cplusplus
}
#endif

#endif /* PMLGPU_FLINT_H */
