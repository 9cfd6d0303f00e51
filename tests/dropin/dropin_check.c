/*
 * Parity check of the FLINT-shaped drop-in (include/pmlgpu_flint.h) against
 * the CPU oracle, written the way a PML user calls the library: FLINT
 * polynomial matrices in, FLINT polynomial matrices out.
 *
 * Follows the reference's own tests of these entry points
 * (flint-extras/src/nmod_poly_mat_extra/test/t-mul_geometric.c,
 * t-mul_waksman.c, t-mulmid.c, t-pmbasis.c, t-mbasis.c: random matrices,
 * equality with a second method, is_approximant_basis on the output) with the
 * oracle as the second method.
 *
 * Prints "ok <case>" per passing case; exits 1 at the first mismatch.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pml_oracle.h"
#include "pmlgpu_flint.h"

static const uint64_t P60 = 1152921504556515329ULL;
static const uint64_t P30 = 998244353ULL;
static const uint64_t P_NOFFT = 723501827743829161ULL; /* no 2^k roots: multi-prime path */
static const uint64_t P63 = 9223372036854775783ULL;     /* largest prime below 2^63 */
static const uint64_t P64 = 18446744073709551557ULL;    /* largest prime below 2^64 */
static uint64_t g_state = 12345;

static void fail(const char *what)
{
    printf("FAIL %s\n", what);
    fflush(stdout);
    exit(1);
}

/* random matrix; entry lengths uniform in [0, maxlen] (ragged) or all maxlen */
static void rand_mat(nmod_poly_mat_t M, slong maxlen, int ragged, int zero_frac)
{
    for (slong i = 0; i < M->r; i++)
        for (slong j = 0; j < M->c; j++) {
            nmod_poly_struct *e = nmod_poly_mat_entry(M, i, j);
            slong len = ragged ? (slong)(orc_splitmix64(&g_state) % (uint64_t)(maxlen + 1)) : maxlen;
            if (zero_frac && orc_splitmix64(&g_state) % 4 == 0)
                len = 0;
            e->length = 0;
            if (len == 0)
                continue;
            nmod_poly_fit_length(e, len);
            orc_rand_uniform(e->coeffs, len, M->modulus, orc_splitmix64(&g_state));
            if (e->coeffs[len - 1] == 0)
                e->coeffs[len - 1] = 1;
            e->length = len;
        }
}

static uint64_t *pack(const nmod_poly_mat_t M, slong len)
{
    uint64_t *b = calloc((size_t)(M->r * M->c * len + 1), sizeof(uint64_t));
    for (slong i = 0; i < M->r; i++)
        for (slong j = 0; j < M->c; j++) {
            const nmod_poly_struct *e = nmod_poly_mat_entry(M, i, j);
            slong l = e->length < len ? e->length : len;
            if (l > 0)
                memcpy(b + (i * M->c + j) * len, e->coeffs, (size_t)l * sizeof(uint64_t));
        }
    return b;
}

/* entries of M equal the oracle's buffer (normalised lengths) */
static void expect_equal(const nmod_poly_mat_t M, const uint64_t *buf, slong len, const char *what)
{
    for (slong i = 0; i < M->r; i++)
        for (slong j = 0; j < M->c; j++) {
            const uint64_t *s = buf + (i * M->c + j) * len;
            slong l = len;
            while (l > 0 && s[l - 1] == 0)
                l--;
            const nmod_poly_struct *e = nmod_poly_mat_entry(M, i, j);
            if (e->length != l) {
                printf("entry (%ld,%ld): length %ld, oracle %ld\n", (long)i, (long)j,
                       (long)e->length, (long)l);
                fail(what);
            }
            if (l && memcmp(e->coeffs, s, (size_t)l * sizeof(uint64_t))) {
                printf("entry (%ld,%ld) differs\n", (long)i, (long)j);
                fail(what);
            }
        }
}

typedef void (*mul_fn)(nmod_poly_mat_t, const nmod_poly_mat_t, const nmod_poly_mat_t);

static void check_mul(mul_fn f, const char *name, slong m, slong k, slong n, slong len, int ragged,
                      int zeros, uint64_t p)
{
    nmod_poly_mat_t A, B, C;
    nmod_poly_mat_init(A, m, k, p);
    nmod_poly_mat_init(B, k, n, p);
    nmod_poly_mat_init(C, m, n, p);
    rand_mat(A, len, ragged, zeros);
    rand_mat(B, len, ragged, zeros);
    /* stale content in the output must be overwritten */
    rand_mat(C, 7, 1, 0);
    f(C, A, B);
    const slong la = nmod_poly_mat_max_length(A), lb = nmod_poly_mat_max_length(B);
    const slong lc = (la && lb) ? la + lb - 1 : 1;
    uint64_t *a = pack(A, la ? la : 1), *b = pack(B, lb ? lb : 1);
    uint64_t *c = calloc((size_t)(m * n * lc + 1), sizeof(uint64_t));
    if (la && lb)
        orc_polmat_mul(c, a, b, m, k, n, la, lb, p);
    expect_equal(C, c, lc, name);
    printf("ok %s %ldx%ldx%ld len<=%ld ragged=%d zeros=%d p=%llu\n", name, (long)m, (long)k,
           (long)n, (long)len, ragged, zeros, (unsigned long long)p);
    free(a);
    free(b);
    free(c);
    nmod_poly_mat_clear(A);
    nmod_poly_mat_clear(B);
    nmod_poly_mat_clear(C);
}

static void check_mul_alias(slong d, slong len, uint64_t p)
{
    nmod_poly_mat_t A, B;
    nmod_poly_mat_init(A, d, d, p);
    nmod_poly_mat_init(B, d, d, p);
    rand_mat(A, len, 1, 0);
    rand_mat(B, len, 1, 0);
    const slong la = nmod_poly_mat_max_length(A), lb = nmod_poly_mat_max_length(B);
    uint64_t *a = pack(A, la), *b = pack(B, lb);
    uint64_t *c = calloc((size_t)(d * d * (la + lb - 1)), sizeof(uint64_t));
    orc_polmat_mul(c, a, b, d, d, d, la, lb, p);
    nmod_poly_mat_multiply(A, A, B); /* res aliases the left operand */
    expect_equal(A, c, la + lb - 1, "multiply aliasing res == A");
    uint64_t *a2 = pack(A, la + lb - 1);
    uint64_t *c2 = calloc((size_t)(d * d * (2 * la + 2 * lb - 3)), sizeof(uint64_t));
    orc_polmat_mul(c2, a2, a2, d, d, d, la + lb - 1, la + lb - 1, p);
    nmod_poly_mat_multiply(A, A, A); /* both operands alias the output */
    expect_equal(A, c2, 2 * la + 2 * lb - 3, "multiply aliasing res == A == B");
    printf("ok multiply aliasing %ldx%ld len<=%ld\n", (long)d, (long)d, (long)len);
    free(a);
    free(b);
    free(c);
    free(a2);
    free(c2);
    nmod_poly_mat_clear(A);
    nmod_poly_mat_clear(B);
}

static void check_mulmid(slong m, slong k, slong n, slong lenA, slong lenB, slong nlo, slong nhi,
                         int underscore, uint64_t p)
{
    nmod_poly_mat_t A, B, R;
    nmod_poly_mat_init(A, m, k, p);
    nmod_poly_mat_init(B, k, n, p);
    nmod_poly_mat_init(R, m, n, p);
    rand_mat(A, lenA, 1, 1);
    rand_mat(B, lenB, 1, 1);
    rand_mat(R, 5, 1, 0);
    const slong la = nmod_poly_mat_max_length(A), lb = nmod_poly_mat_max_length(B);
    if (underscore)
        _nmod_poly_mat_mulmid(R, A, la, B, lb, nlo, nhi);
    else
        nmod_poly_mat_mulmid(R, A, B, nlo, nhi);
    const slong lr = nhi > nlo ? nhi - nlo : 1;
    uint64_t *a = pack(A, la ? la : 1), *b = pack(B, lb ? lb : 1);
    uint64_t *r = calloc((size_t)(m * n * lr + 1), sizeof(uint64_t));
    if (la && lb && nhi > nlo)
        orc_polmat_mulmid(r, a, b, m, k, n, la, lb, nlo, nhi, p);
    expect_equal(R, r, lr, underscore ? "_nmod_poly_mat_mulmid" : "nmod_poly_mat_mulmid");
    printf("ok %smulmid %ldx%ldx%ld len %ld,%ld [%ld,%ld) p=%llu\n", underscore ? "_" : "",
           (long)m, (long)k, (long)n, (long)lenA, (long)lenB, (long)nlo, (long)nhi,
           (unsigned long long)p);
    free(a);
    free(b);
    free(r);
    nmod_poly_mat_clear(A);
    nmod_poly_mat_clear(B);
    nmod_poly_mat_clear(R);
}

static void check_basis(int iterative, slong m, slong n, slong lenF, slong order, slong shift_step,
                        uint64_t p)
{
    const char *name = iterative ? "nmod_poly_mat_mbasis" : "nmod_poly_mat_pmbasis";
    nmod_poly_mat_t F, P;
    nmod_poly_mat_init(F, m, n, p);
    nmod_poly_mat_init(P, m, m, p);
    rand_mat(F, lenF, 0, 0);
    slong *shift = malloc((size_t)m * sizeof(slong)), *shift0 = malloc((size_t)m * sizeof(slong));
    int64_t *oshift = malloc((size_t)m * sizeof(int64_t));
    for (slong i = 0; i < m; i++)
        shift[i] = shift0[i] = oshift[i] = (shift_step * i) % 7;
    if (iterative)
        nmod_poly_mat_mbasis(P, shift, F, (ulong)order);
    else
        nmod_poly_mat_pmbasis(P, shift, F, order);
    uint64_t *f = pack(F, lenF);
    uint64_t *OP = NULL;
    int64_t olen = 0;
    if (iterative)
        orc_mbasis(&OP, &olen, oshift, f, m, n, lenF, order, p);
    else
        orc_pmbasis(&OP, &olen, oshift, f, m, n, lenF, order, p);
    expect_equal(P, OP, olen, name);
    for (slong i = 0; i < m; i++)
        if (shift[i] != oshift[i])
            fail("shifted row degree differs from the oracle");
    if (!nmod_poly_mat_is_approximant_basis(P, F, order, shift0, ROW_LOWER))
        fail("is_approximant_basis rejects the computed basis");
    /* the oracle's certificate agrees */
    uint64_t *pp = pack(P, olen);
    if (orc_is_approximant_basis(pp, olen, f, m, n, lenF, order, shift0, p, 7) != 0)
        fail("oracle certificate rejects the computed basis");
    printf("ok %s %ldx%ld order %ld p=%llu\n", name, (long)m, (long)n, (long)order,
           (unsigned long long)p);

    /* x * (row 0): still approximants, still weak Popov with the same pivots
     * and det = lambda x^(D+1) with D+1 the new diagonal degree sum, but
     * [P(0) | C] loses rank: only the generation test catches it */
    nmod_poly_struct *e;
    for (slong j = 0; j < m; j++) {
        e = nmod_poly_mat_entry(P, 0, j);
        if (e->length == 0)
            continue;
        nmod_poly_fit_length(e, e->length + 1);
        memmove(e->coeffs + 1, e->coeffs, (size_t)e->length * sizeof(uint64_t));
        e->coeffs[0] = 0;
        e->length++;
    }
    if (nmod_poly_mat_is_approximant_basis(P, F, order, shift0, ROW_LOWER))
        fail("is_approximant_basis accepts x * row 0");
    /* (x + 1) * row 0 instead: approximants, weak Popov, deg det = D and
     * [P(0) | C] unchanged, so the reference's certificate, whose determinant
     * test checks the degree only (verification.c:73-84), accepts it */
    for (slong j = 0; j < m; j++) {
        e = nmod_poly_mat_entry(P, 0, j);
        for (slong d = 0; d + 1 < e->length; d++) {
            /* residues of up to 64 bits: the sum may carry out */
            const uint64_t a = e->coeffs[d], b = e->coeffs[d + 1], t = a + b;
            e->coeffs[d] = (t < a || t >= p) ? t - p : t;
        }
    }
    if (!nmod_poly_mat_is_approximant_basis(P, F, order, shift0, ROW_LOWER))
        fail("is_approximant_basis rejects (x + 1) * row 0, which the reference accepts");
    if (m >= 2) {
        /* rows 0 and 1 exchanged: no longer ordered weak Popov */
        for (slong j = 0; j < m; j++) {
            nmod_poly_struct t = *nmod_poly_mat_entry(P, 0, j);
            *nmod_poly_mat_entry(P, 0, j) = *nmod_poly_mat_entry(P, 1, j);
            *nmod_poly_mat_entry(P, 1, j) = t;
        }
        if (nmod_poly_mat_is_approximant_basis(P, F, order, shift0, ROW_LOWER))
            fail("is_approximant_basis accepts a basis with two rows exchanged");
        for (slong j = 0; j < m; j++) {
            nmod_poly_struct t = *nmod_poly_mat_entry(P, 0, j);
            *nmod_poly_mat_entry(P, 0, j) = *nmod_poly_mat_entry(P, 1, j);
            *nmod_poly_mat_entry(P, 1, j) = t;
        }
    }
    /* a changed low coefficient breaks the residual */
    e = nmod_poly_mat_entry(P, m - 1, m - 1);
    if (e->length == 0) {
        nmod_poly_fit_length(e, 1);
        e->coeffs[0] = 0;
        e->length = 1;
    }
    e->coeffs[0] = (e->coeffs[0] + 1) % p;
    while (e->length > 0 && e->coeffs[e->length - 1] == 0)
        e->length--;
    if (nmod_poly_mat_is_approximant_basis(P, F, order, shift0, ROW_LOWER))
        fail("is_approximant_basis accepts a perturbed basis");
    printf("ok is_approximant_basis rejects perturbed bases %ldx%ld\n", (long)m, (long)n);
    orc_free(OP);
    free(f);
    free(pp);
    free(shift);
    free(shift0);
    free(oshift);
    nmod_poly_mat_clear(F);
    nmod_poly_mat_clear(P);
}

/* shifted degree and pivot (rightmost for ROW_LOWER / COL_UPPER, leftmost
 * for ROW_UPPER / COL_LOWER) of row or column t, as the reference's
 * nmod_poly_mat_row_degree / _column_degree and _pivot_index */
static void line_profile(const nmod_poly_mat_t K, slong t, const slong *shift, orientation_t o,
                         slong *deg, slong *piv)
{
    const int rows = o == ROW_LOWER || o == ROW_UPPER, last = o == ROW_LOWER || o == COL_UPPER;
    const slong len = rows ? K->c : K->r;
    slong mx = 0, pv = -1;
    for (slong j = 0; j < len; j++) {
        const nmod_poly_struct *e = rows ? nmod_poly_mat_entry(K, t, j) : nmod_poly_mat_entry(K, j, t);
        if (e->length == 0)
            continue;
        const slong v = e->length - 1 + shift[j];
        if (pv == -1 || (last ? v >= mx : v > mx)) {
            pv = j;
            mx = v;
        }
    }
    *deg = mx;
    *piv = pv;
}

/* the first `nul` rows (left) or columns (right) of ker as their own matrix */
static void basis_part(nmod_poly_mat_t out, const nmod_poly_mat_t ker, slong nul, int left)
{
    nmod_poly_mat_init(out, left ? nul : ker->r, left ? ker->c : nul, ker->modulus);
    for (slong i = 0; i < out->r; i++)
        for (slong j = 0; j < out->c; j++) {
            const nmod_poly_struct *e = nmod_poly_mat_entry(ker, i, j);
            nmod_poly_struct *d = nmod_poly_mat_entry(out, i, j);
            if (e->length) {
                nmod_poly_fit_length(d, e->length);
                memcpy(d->coeffs, e->coeffs, (size_t)e->length * sizeof(uint64_t));
            }
            d->length = e->length;
        }
}

/* Kernel bases through the drop-in, following the reference's kernel test
 * (t-kernel.c:28-260): for F = G * H of rank r (m x n), every orientation of
 * nmod_poly_mat_kernel and both left-kernel algorithms give a basis that
 * nmod_poly_mat_is_kernel accepts, of nullity dim - r, whose shifted degrees
 * and pivot indices are the returned ones; the two algorithms agree on them
 * (they depend only on the kernel module and the shift); a perturbed basis
 * is rejected. */
static void check_kernel(slong m, slong n, slong len, slong r, uint64_t p)
{
    nmod_poly_mat_t G, H, F;
    nmod_poly_mat_init(G, m, r, p);
    nmod_poly_mat_init(H, r, n, p);
    nmod_poly_mat_init(F, m, n, p);
    rand_mat(G, (len + 1) / 2, 0, 0);
    rand_mat(H, len - (len + 1) / 2 + 1, 0, 0);
    nmod_poly_mat_multiply(F, G, H);
    const orientation_t ors[4] = {ROW_LOWER, ROW_UPPER, COL_LOWER, COL_UPPER};
    slong piv_first[2][128], deg_first[2][128], nul_first[2] = {-1, -1};
    for (int a = 0; a < 6; a++) {
        const orientation_t o = a < 4 ? ors[a] : ROW_LOWER;
        const int left = o == ROW_LOWER || o == ROW_UPPER;
        const slong dim = left ? m : n;
        nmod_poly_mat_t ker;
        nmod_poly_mat_init(ker, dim, dim, p);
        slong shift[128], shift0[128], piv[128];
        for (slong i = 0; i < dim; i++)
            shift[i] = shift0[i] = (3 * i + (a < 4 ? a : 4)) % 5; /* cases 4, 5: one shift */
        slong nul;
        const char *name;
        if (a < 4) {
            nul = nmod_poly_mat_kernel(ker, piv, shift, F, ORD_WEAK_POPOV, o);
            name = "nmod_poly_mat_kernel";
        } else if (a == 4) {
            nul = nmod_poly_mat_kernel_via_approx(ker, piv, shift, F);
            name = "nmod_poly_mat_kernel_via_approx";
        } else {
            nul = nmod_poly_mat_kernel_zls_approx(ker, piv, shift, F);
            name = "nmod_poly_mat_kernel_zls_approx";
        }
        if (nul != dim - r)
            fail("kernel: nullity is not dim - rank");
        nmod_poly_mat_t kz;
        basis_part(kz, ker, nul, left);
        if (!nmod_poly_mat_is_kernel(kz, shift0, F, ORD_WEAK_POPOV, o))
            fail("kernel: is_kernel rejects the computed basis");
        for (slong t = 0; t < nul; t++) {
            slong d, pv;
            line_profile(kz, t, shift0, o, &d, &pv);
            if (d != shift[t] || pv != piv[t])
                fail("kernel: returned degrees / pivots differ from the basis");
        }
        if (a >= 4) {
            const int w = a - 4;
            nul_first[w] = nul;
            for (slong t = 0; t < nul; t++) {
                piv_first[w][t] = piv[t];
                deg_first[w][t] = shift[t];
            }
        }
        if (nul > 0 && a == 0) {
            /* a changed low coefficient leaves the kernel */
            nmod_poly_struct *e = nmod_poly_mat_entry(kz, 0, piv[0]);
            e->coeffs[0] = (e->coeffs[0] + 1) % p;
            while (e->length > 0 && e->coeffs[e->length - 1] == 0)
                e->length--;
            if (nmod_poly_mat_is_kernel(kz, shift0, F, ORD_WEAK_POPOV, o))
                fail("kernel: is_kernel accepts a perturbed basis");
        }
        printf("ok %s orient %d %ldx%ld rank %ld nullity %ld p=%llu\n", name, (int)o, (long)m,
               (long)n, (long)r, (long)nul, (unsigned long long)p);
        nmod_poly_mat_clear(kz);
        nmod_poly_mat_clear(ker);
    }
    if (nul_first[0] != nul_first[1])
        fail("kernel: the two algorithms disagree on the nullity");
    for (slong t = 0; t < nul_first[0]; t++)
        if (piv_first[0][t] != piv_first[1][t] || deg_first[0][t] != deg_first[1][t])
            fail("kernel: the two algorithms disagree on the pivot profile");
    nmod_poly_mat_clear(G);
    nmod_poly_mat_clear(H);
    nmod_poly_mat_clear(F);
}

/* More distinct moduli than the drop-in caches contexts for (32), as the
 * reference's tests draw a new random prime per iteration
 * (t-mul_geometric.c:59-60, t-pmbasis.c:468-476): each product is still
 * exact, and the cache stays at its bound. */
static void check_many_moduli(void)
{
    int done = 0;
    for (int i = 0; i < 40; i++) {
        uint64_t p = ((uint64_t)1 << (12 + i)) + 1000003ULL * (uint64_t)i;
        while (!orc_is_prime(p))
            p++;
        nmod_poly_mat_t A, B, C;
        nmod_poly_mat_init(A, 2, 3, p);
        nmod_poly_mat_init(B, 3, 2, p);
        nmod_poly_mat_init(C, 2, 2, p);
        rand_mat(A, 20, 1, 0);
        rand_mat(B, 20, 1, 0);
        nmod_poly_mat_multiply(C, A, B);
        const slong la = nmod_poly_mat_max_length(A), lb = nmod_poly_mat_max_length(B);
        const slong lc = (la && lb) ? la + lb - 1 : 1;
        uint64_t *a = pack(A, la ? la : 1), *b = pack(B, lb ? lb : 1);
        uint64_t *c = calloc((size_t)(4 * lc + 1), sizeof(uint64_t));
        if (la && lb)
            orc_polmat_mul(c, a, b, 2, 3, 2, la, lb, p);
        expect_equal(C, c, lc, "multiply over many moduli");
        free(a);
        free(b);
        free(c);
        nmod_poly_mat_clear(A);
        nmod_poly_mat_clear(B);
        nmod_poly_mat_clear(C);
        done++;
    }
    if (pmlg_dropin_cached_contexts() > 32)
        fail("the drop-in keeps more than 32 contexts");
    printf("ok multiply over %d distinct moduli (%d contexts cached)\n", done,
           pmlg_dropin_cached_contexts());
}

int main(void)
{
    check_mul(nmod_poly_mat_multiply, "nmod_poly_mat_multiply", 5, 7, 3, 300, 1, 1, P60);
    check_mul(nmod_poly_mat_multiply, "nmod_poly_mat_multiply", 16, 16, 16, 1000, 0, 0, P60);
    check_mul(nmod_poly_mat_multiply, "nmod_poly_mat_multiply", 9, 4, 11, 513, 1, 0, P30);
    check_mul(nmod_poly_mat_multiply, "nmod_poly_mat_multiply", 1, 6, 1, 64, 1, 1, P30);
    check_mul(nmod_poly_mat_multiply, "nmod_poly_mat_multiply", 3, 3, 3, 1, 0, 0, P60);
    check_mul(nmod_poly_mat_mul_geometric, "nmod_poly_mat_mul_geometric", 6, 5, 4, 200, 1, 0, P60);
    check_mul(nmod_poly_mat_mul_vandermonde1, "nmod_poly_mat_mul_vandermonde1", 6, 5, 4, 200, 1, 0, P60);
    check_mul(nmod_poly_mat_mul_vandermonde2, "nmod_poly_mat_mul_vandermonde2", 6, 5, 4, 200, 1, 0, P30);
    check_mul(nmod_poly_mat_mul_waksman, "nmod_poly_mat_mul_waksman", 6, 5, 4, 200, 1, 0, P30);
    check_mul(nmod_poly_mat_multiply, "nmod_poly_mat_multiply", 6, 5, 7, 400, 1, 1, P_NOFFT);
    /* 63- and 64-bit moduli (the reference's tests draw primes up to 64 bits) */
    check_mul(nmod_poly_mat_multiply, "nmod_poly_mat_multiply", 4, 5, 3, 60, 1, 1, P63);
    check_mul(nmod_poly_mat_multiply, "nmod_poly_mat_multiply", 3, 4, 5, 50, 1, 0, P64);
    /* a product of length 199,999 (transform length 2^18) */
    check_mul(nmod_poly_mat_multiply, "nmod_poly_mat_multiply", 2, 2, 2, 100000, 0, 0, P60);
    check_mul_alias(6, 100, P60);
    check_mulmid(7, 5, 6, 200, 300, 100, 300, 0, P60);
    check_mulmid(7, 5, 6, 200, 300, 0, 1000, 0, P60); /* nhi beyond the product length */
    check_mulmid(4, 4, 4, 900, 50, 10, 60, 0, P30);   /* left operand longer than nhi */
    check_mulmid(4, 4, 4, 50, 60, 80, 80, 0, P30);    /* empty range */
    check_mulmid(8, 8, 4, 65, 128, 64, 128, 1, P30);  /* PM-Basis residual shape */
    check_mulmid(3, 2, 5, 40, 400, 39, 400, 1, P60);
    check_mulmid(3, 4, 2, 50, 70, 20, 90, 0, P64);
    check_basis(0, 6, 3, 100, 100, 3, P60);
    check_basis(0, 8, 4, 150, 150, 0, P30);
    check_basis(0, 5, 5, 64, 64, 1, P30); /* square F */
    check_basis(1, 4, 2, 40, 40, 2, P30);
    check_basis(1, 7, 3, 25, 25, 0, P60);
    check_basis(0, 5, 2, 90, 90, 1, P_NOFFT);
    check_basis(0, 5, 3, 60, 60, 1, P64);   /* 64-bit prime: plain-residue M-Basis steps */
    check_basis(1, 4, 2, 30, 30, 2, P63);
    check_basis(0, 160, 80, 40, 40, 1, P60); /* more than 128 rows */
    check_basis(0, 5, 2, 120, 120, 1, 2);    /* p = 2 (a random 2-bit prime in the reference's tests) */
    check_basis(1, 4, 3, 30, 30, 0, 3);
    check_kernel(6, 2, 8, 2, P60);
    check_kernel(7, 4, 5, 2, P30);
    check_kernel(5, 5, 6, 3, P30);
    check_kernel(6, 2, 7, 1, 2);
    check_kernel(9, 3, 40, 3, P60);
    check_kernel(4, 6, 3, 1, P_NOFFT);
    check_kernel(6, 2, 8, 2, P64);
    check_many_moduli();
    printf("all drop-in checks passed\n");
    return 0;
}
