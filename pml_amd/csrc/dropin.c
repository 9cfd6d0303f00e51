/*
 * FLINT-shaped drop-in entry points of libpmlgpu (declared in
 * include/pmlgpu_flint.h): PML's nmod_poly_mat_multiply / mulmid /
 * pmbasis / mbasis / is_approximant_basis, computed on the GPU.
 *
 * Each call packs the FLINT polynomial matrices into entry-major buffers
 * (entry (i,j) coefficient d at [(i*cols + j)*len + d]), runs the device
 * routine on a context for the matrices' modulus, and unpacks the result into
 * the caller's matrix, growing and normalising its entries.  Contexts are
 * cached per modulus for the life of the process; a context serialises its own
 * calls, so several threads may use the drop-in at once.  The device is
 * PMLG_DEVICE (default 0).  Failures abort with a message, as flint_throw
 * does: the drop-in has no CPU fallback.
 *
 * Entry growth uses FLINT's nmod_poly_fit_length when the process has it
 * (the drop-in then lives beside a real FLINT, whose memory functions own the
 * entries), else the pmlg_nmod_poly_* functions defined here.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/pmlgpu_flint.h"

#ifndef PMLG_USE_FLINT
#include <stddef.h>
/* the FLINT >= 3.3 layouts a FLINT 3.4 caller hands over (x86-64 LP64) */
_Static_assert(sizeof(nmod_t) == 24, "nmod_t is three words");
_Static_assert(sizeof(nmod_poly_struct) == 48 && offsetof(nmod_poly_struct, coeffs) == 0 &&
                   offsetof(nmod_poly_struct, alloc) == 8 &&
                   offsetof(nmod_poly_struct, length) == 16 &&
                   offsetof(nmod_poly_struct, mod) == 24,
               "nmod_poly_struct layout");
_Static_assert(sizeof(nmod_poly_mat_struct) == 32 && offsetof(nmod_poly_mat_struct, entries) == 0 &&
                   offsetof(nmod_poly_mat_struct, r) == 8 &&
                   offsetof(nmod_poly_mat_struct, c) == 16 &&
                   offsetof(nmod_poly_mat_struct, modulus) == 24,
               "nmod_poly_mat_struct layout (FLINT >= 3.3: no rows array)");
#endif

/* ------------------------------------------------------------------------ */
/* memory management of the FLINT subset                                     */
/* ------------------------------------------------------------------------ */

static void nmod_init_fields(nmod_t *mod, ulong n)
{
    /* FLINT's nmod_t: modulus, normalisation shift and the 2-by-1 reciprocal
     * of the normalised modulus, floor((2^128 - 1) / (n << norm)) - 2^64. */
    mod->n = n;
    mod->norm = n ? (ulong)__builtin_clzl(n) : 0;
    if (n) {
        const unsigned __int128 d = (unsigned __int128)(n << mod->norm);
        mod->ninv = (ulong)((~(unsigned __int128)0) / d - ((unsigned __int128)1 << 64));
    } else {
        mod->ninv = 0;
    }
}

void pmlg_nmod_poly_init(nmod_poly_t poly, ulong n)
{
    poly->coeffs = NULL;
    poly->alloc = 0;
    poly->length = 0;
    nmod_init_fields(&poly->mod, n);
}

void pmlg_nmod_poly_clear(nmod_poly_t poly)
{
    free(poly->coeffs);
    poly->coeffs = NULL;
    poly->alloc = 0;
    poly->length = 0;
}

void pmlg_nmod_poly_fit_length(nmod_poly_t poly, slong len)
{
    if (len <= poly->alloc)
        return;
    slong want = len > 2 * poly->alloc ? len : 2 * poly->alloc;
    ulong *c = (ulong *)realloc(poly->coeffs, (size_t)want * sizeof(ulong));
    if (!c) {
        fprintf(stderr, "pmlgpu: out of host memory (%ld coefficients)\n", (long)want);
        abort();
    }
    poly->coeffs = c;
    poly->alloc = want;
}

static void poly_normalise(nmod_poly_struct *poly)
{
    while (poly->length > 0 && poly->coeffs[poly->length - 1] == 0)
        poly->length--;
}

void pmlg_nmod_poly_set_coeff_ui(nmod_poly_t poly, slong j, ulong c)
{
    c %= poly->mod.n;
    if (j >= poly->length) {
        if (c == 0)
            return;
        pmlg_nmod_poly_fit_length(poly, j + 1);
        memset(poly->coeffs + poly->length, 0, (size_t)(j - poly->length) * sizeof(ulong));
        poly->length = j + 1;
    }
    poly->coeffs[j] = c;
    if (j == poly->length - 1)
        poly_normalise(poly);
}

ulong pmlg_nmod_poly_get_coeff_ui(const nmod_poly_t poly, slong j)
{
    return (j >= 0 && j < poly->length) ? poly->coeffs[j] : 0;
}

void pmlg_nmod_poly_mat_init(nmod_poly_mat_t mat, slong rows, slong cols, ulong n)
{
    mat->entries = NULL;
    mat->r = rows;
    mat->c = cols;
    mat->modulus = n;
    if (rows > 0 && cols > 0) {
        mat->entries = (nmod_poly_struct *)malloc((size_t)(rows * cols) * sizeof(nmod_poly_struct));
        if (!mat->entries)
            abort();
        for (slong e = 0; e < rows * cols; e++)
            pmlg_nmod_poly_init(mat->entries + e, n);
    }
}

void pmlg_nmod_poly_mat_clear(nmod_poly_mat_t mat)
{
    if (mat->entries)
        for (slong e = 0; e < mat->r * mat->c; e++)
            pmlg_nmod_poly_clear(mat->entries + e);
    free(mat->entries);
    mat->entries = NULL;
}

void pmlg_nmod_poly_mat_zero(nmod_poly_mat_t mat)
{
    for (slong i = 0; i < mat->r; i++)
        for (slong j = 0; j < mat->c; j++)
            nmod_poly_mat_entry(mat, i, j)->length = 0;
}

slong pmlg_nmod_poly_mat_max_length(const nmod_poly_mat_t mat)
{
    slong len = 0;
    for (slong i = 0; i < mat->r; i++)
        for (slong j = 0; j < mat->c; j++) {
            const slong l = nmod_poly_mat_entry(mat, i, j)->length;
            len = l > len ? l : len;
        }
    return len;
}

int pmlg_nmod_poly_mat_equal(const nmod_poly_mat_t a, const nmod_poly_mat_t b)
{
    if (a->r != b->r || a->c != b->c)
        return 0;
    for (slong i = 0; i < a->r; i++)
        for (slong j = 0; j < a->c; j++) {
            const nmod_poly_struct *x = nmod_poly_mat_entry(a, i, j);
            const nmod_poly_struct *y = nmod_poly_mat_entry(b, i, j);
            if (x->length != y->length ||
                (x->length && memcmp(x->coeffs, y->coeffs, (size_t)x->length * sizeof(ulong))))
                return 0;
        }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* contexts, errors, packing                                                 */
/* ------------------------------------------------------------------------ */

#define CTX_CACHE 32

/* Contexts cached per modulus, least recently used replaced first.  A call
 * holds its context from ctx_acquire to ctx_release, so an entry in use is
 * never replaced; when all CTX_CACHE entries are in use at once, the call
 * gets a context of its own, destroyed when it is released. */
static pthread_mutex_t g_cache_mu = PTHREAD_MUTEX_INITIALIZER;
static struct {
    uint64_t p;
    pmlg_ctx ctx;
    uint64_t stamp; /* last acquisition (LRU order) */
    int users;      /* calls holding the context now */
} g_cache[CTX_CACHE];
static int g_cache_n = 0;
static uint64_t g_cache_clock = 0;
static int g_cache_atexit = 0;

typedef void (*fit_length_fn)(nmod_poly_struct *, slong);
static fit_length_fn g_fit = NULL;
static pthread_once_t g_fit_once = PTHREAD_ONCE_INIT;

static void resolve_fit(void)
{
    /* a FLINT loaded in the process owns the entries' memory */
    void *f = dlsym(RTLD_DEFAULT, "nmod_poly_fit_length");
    g_fit = f ? (fit_length_fn)f : (fit_length_fn)pmlg_nmod_poly_fit_length;
}

static void die(const char *fn, const char *what)
{
    fprintf(stderr, "Exception (%s). %s\n", fn, what);
    fflush(stderr);
    abort();
}

static void check(pmlg_status st, const char *fn)
{
    if (st != PMLG_OK)
        die(fn, pmlg_last_error());
}

static void destroy_cache(void)
{
    for (int i = 0; i < g_cache_n; i++)
        pmlg_ctx_destroy(g_cache[i].ctx);
    g_cache_n = 0;
}

static pmlg_ctx new_ctx(uint64_t p, const char *fn)
{
    const char *dev = getenv("PMLG_DEVICE");
    pmlg_ctx c = NULL;
    if (pmlg_ctx_create(&c, p, dev ? atoi(dev) : 0) != PMLG_OK)
        die(fn, pmlg_last_error());
    return c;
}

static pmlg_ctx ctx_acquire(uint64_t p, const char *fn)
{
    pthread_mutex_lock(&g_cache_mu);
    const uint64_t now = ++g_cache_clock;
    for (int i = 0; i < g_cache_n; i++)
        if (g_cache[i].p == p) {
            g_cache[i].users++;
            g_cache[i].stamp = now;
            pmlg_ctx c = g_cache[i].ctx;
            pthread_mutex_unlock(&g_cache_mu);
            return c;
        }
    int slot = -1;
    if (g_cache_n < CTX_CACHE) {
        slot = g_cache_n;
    } else {
        /* least recently used idle entry; its destruction synchronises its
         * stream, so work still queued there completes first */
        for (int i = 0; i < g_cache_n; i++)
            if (g_cache[i].users == 0 && (slot < 0 || g_cache[i].stamp < g_cache[slot].stamp))
                slot = i;
        if (slot < 0) {
            pthread_mutex_unlock(&g_cache_mu);
            return new_ctx(p, fn); /* every entry in use: uncached */
        }
        pmlg_ctx_destroy(g_cache[slot].ctx);
        g_cache[slot].ctx = NULL;
    }
    pmlg_ctx c = NULL;
    const char *dev = getenv("PMLG_DEVICE");
    if (pmlg_ctx_create(&c, p, dev ? atoi(dev) : 0) != PMLG_OK) {
        if (slot < g_cache_n) { /* the replaced entry is gone: close the gap */
            g_cache[slot] = g_cache[g_cache_n - 1];
            g_cache_n--;
        }
        pthread_mutex_unlock(&g_cache_mu);
        die(fn, pmlg_last_error());
    }
    if (!g_cache_atexit) {
        atexit(destroy_cache);
        g_cache_atexit = 1;
    }
    g_cache[slot].p = p;
    g_cache[slot].ctx = c;
    g_cache[slot].stamp = now;
    g_cache[slot].users = 1;
    if (slot == g_cache_n)
        g_cache_n++;
    pthread_mutex_unlock(&g_cache_mu);
    return c;
}

static void ctx_release(pmlg_ctx c)
{
    pthread_mutex_lock(&g_cache_mu);
    for (int i = 0; i < g_cache_n; i++)
        if (g_cache[i].ctx == c) {
            g_cache[i].users--;
            pthread_mutex_unlock(&g_cache_mu);
            return;
        }
    pthread_mutex_unlock(&g_cache_mu);
    pmlg_ctx_destroy(c); /* an uncached context */
}

/* number of cached contexts (for the drop-in checks) */
int pmlg_dropin_cached_contexts(void)
{
    pthread_mutex_lock(&g_cache_mu);
    const int n = g_cache_n;
    pthread_mutex_unlock(&g_cache_mu);
    return n;
}

static uint64_t *xalloc(size_t words, const char *fn)
{
    uint64_t *b = (uint64_t *)malloc((words ? words : 1) * sizeof(uint64_t));
    if (!b)
        die(fn, "out of host memory");
    return b;
}

/* entry-major copy of the first `len` coefficients of each entry, zero padded */
static uint64_t *pack(const nmod_poly_mat_struct *M, slong len, const char *fn)
{
    const slong r = M->r, c = M->c;
    uint64_t *b = xalloc((size_t)(r * c * len), fn);
    for (slong i = 0; i < r; i++)
        for (slong j = 0; j < c; j++) {
            const nmod_poly_struct *e = nmod_poly_mat_entry(M, i, j);
            uint64_t *d = b + (i * c + j) * len;
            const slong l = e->length < len ? e->length : len;
            if (l > 0)
                memcpy(d, e->coeffs, (size_t)l * sizeof(uint64_t));
            if (l < len)
                memset(d + (l > 0 ? l : 0), 0, (size_t)(len - (l > 0 ? l : 0)) * sizeof(uint64_t));
        }
    return b;
}

/* res (r x c, entries of padded length len in buf) <- buf, normalised */
static void unpack(nmod_poly_mat_struct *res, const uint64_t *buf, slong len)
{
    pthread_once(&g_fit_once, resolve_fit);
    const slong r = res->r, c = res->c;
    for (slong i = 0; i < r; i++)
        for (slong j = 0; j < c; j++) {
            nmod_poly_struct *e = nmod_poly_mat_entry(res, i, j);
            const uint64_t *s = buf + (i * c + j) * len;
            slong l = len;
            while (l > 0 && s[l - 1] == 0)
                l--;
            if (l > 0) {
                g_fit(e, l);
                memcpy(e->coeffs, s, (size_t)l * sizeof(uint64_t));
            }
            e->length = l;
        }
}

static void need_dims(int ok, const char *fn)
{
    if (!ok)
        die(fn, "incompatible matrix dimensions");
}

/* ------------------------------------------------------------------------ */
/* products                                                                  */
/* ------------------------------------------------------------------------ */

/* do the entry structs of X and Y overlap (res aliasing an operand, also
 * through windows)? */
static int shares_entries(const nmod_poly_mat_struct *X, const nmod_poly_mat_struct *Y)
{
    if (X->r == 0 || X->c == 0 || Y->r == 0 || Y->c == 0)
        return 0;
    uintptr_t xl = UINTPTR_MAX, xh = 0, yl = UINTPTR_MAX, yh = 0;
    for (slong i = 0; i < X->r; i++) {
        const uintptr_t a = (uintptr_t)nmod_poly_mat_entry(X, i, 0);
        const uintptr_t b = (uintptr_t)(nmod_poly_mat_entry(X, i, X->c - 1) + 1);
        xl = a < xl ? a : xl;
        xh = b > xh ? b : xh;
    }
    for (slong i = 0; i < Y->r; i++) {
        const uintptr_t a = (uintptr_t)nmod_poly_mat_entry(Y, i, 0);
        const uintptr_t b = (uintptr_t)(nmod_poly_mat_entry(Y, i, Y->c - 1) + 1);
        yl = a < yl ? a : yl;
        yh = b > yh ? b : yh;
    }
    return xl < yh && yl < xh;
}

/* coefficient pointers and lengths of M's entries, lengths clipped to clip */
static void entry_lists(const nmod_poly_mat_struct *M, slong clip, const uint64_t **ptr,
                        int64_t *len)
{
    for (slong i = 0; i < M->r; i++)
        for (slong j = 0; j < M->c; j++) {
            const nmod_poly_struct *e = nmod_poly_mat_entry(M, i, j);
            const slong l = e->length < clip ? e->length : clip;
            ptr[i * M->c + j] = e->coeffs;
            len[i * M->c + j] = l > 0 ? l : 0;
        }
}

/* res = (A*B mod x^nhi) div x^nlo from the first len1 / len2 coefficients of
 * the entries of A / B, through the entry-wise transfers of the library: the
 * operands' coefficient arrays are read in place (no packed host copy) and
 * the result is written straight into res's entries, which are grown to
 * nhi - nlo words first.  When res shares entries with A or B the result goes
 * through a host buffer first, so that growing res cannot move an operand. */
static void product_entries(nmod_poly_mat_struct *res, const nmod_poly_mat_struct *A, slong len1,
                            const nmod_poly_mat_struct *B, slong len2, slong nlo, slong nhi,
                            const char *fn)
{
    pthread_once(&g_fit_once, resolve_fit);
    const slong m = A->r, k = A->c, n = B->c, w = nhi - nlo;
    const uint64_t **ap = (const uint64_t **)malloc((size_t)(m * k + 1) * sizeof(void *));
    const uint64_t **bp = (const uint64_t **)malloc((size_t)(k * n + 1) * sizeof(void *));
    uint64_t **cp = (uint64_t **)malloc((size_t)(m * n + 1) * sizeof(void *));
    int64_t *al = (int64_t *)malloc((size_t)(m * k + 1) * sizeof(int64_t));
    int64_t *bl = (int64_t *)malloc((size_t)(k * n + 1) * sizeof(int64_t));
    int64_t *cl = (int64_t *)malloc((size_t)(m * n + 1) * sizeof(int64_t));
    if (!ap || !bp || !cp || !al || !bl || !cl)
        die(fn, "out of host memory");
    entry_lists(A, len1, ap, al);
    entry_lists(B, len2, bp, bl);
    const int alias = shares_entries(res, A) || shares_entries(res, B);
    uint64_t *buf = NULL;
    if (alias) {
        buf = xalloc((size_t)(m * n * w), fn);
        for (slong e = 0; e < m * n; e++)
            cp[e] = buf + e * w;
    } else {
        for (slong i = 0; i < m; i++)
            for (slong j = 0; j < n; j++) {
                nmod_poly_struct *e = nmod_poly_mat_entry(res, i, j);
                g_fit(e, w);
                cp[i * n + j] = e->coeffs;
            }
    }
    pmlg_ctx ctx = ctx_acquire(A->modulus, fn);
    check(pmlg_polmat_mulmid_entries(ctx, cp, cl, ap, al, bp, bl, m, k, n, nlo, nhi,
                                     PMLG_METHOD_AUTO),
          fn);
    ctx_release(ctx);
    for (slong i = 0; i < m; i++)
        for (slong j = 0; j < n; j++) {
            nmod_poly_struct *e = nmod_poly_mat_entry(res, i, j);
            const slong l = cl[i * n + j];
            if (alias && l > 0) {
                g_fit(e, l);
                memcpy(e->coeffs, cp[i * n + j], (size_t)l * sizeof(uint64_t));
            }
            e->length = l;
        }
    free(buf);
    free(ap);
    free(bp);
    free(cp);
    free(al);
    free(bl);
    free(cl);
}

/* nmod_poly_mat_multiply (mul.c:19-97): the product is unique, so the
 * reference's choice among its evaluation back-ends has no counterpart here;
 * the device picks the pointwise kernel (int8-limb MFMA or plain) itself. */
static void mul_impl(nmod_poly_mat_struct *res, const nmod_poly_mat_struct *A,
                     const nmod_poly_mat_struct *B, const char *fn)
{
    need_dims(A->c == B->r && res->r == A->r && res->c == B->c, fn);
    const slong lenA = pmlg_nmod_poly_mat_max_length(A);
    const slong lenB = pmlg_nmod_poly_mat_max_length(B);
    if (lenA == 0 || lenB == 0 || A->r == 0 || B->c == 0) {
        pmlg_nmod_poly_mat_zero(res);
        return;
    }
    product_entries(res, A, lenA, B, lenB, 0, lenA + lenB - 1, fn);
}

void nmod_poly_mat_multiply(nmod_poly_mat_t res, const nmod_poly_mat_t A, const nmod_poly_mat_t B)
{
    mul_impl(res, A, B, "nmod_poly_mat_multiply");
}

void nmod_poly_mat_mul_geometric(nmod_poly_mat_t res, const nmod_poly_mat_t A,
                                 const nmod_poly_mat_t B)
{
    mul_impl(res, A, B, "nmod_poly_mat_mul_geometric");
}

void nmod_poly_mat_mul_vandermonde1(nmod_poly_mat_t res, const nmod_poly_mat_t A,
                                    const nmod_poly_mat_t B)
{
    mul_impl(res, A, B, "nmod_poly_mat_mul_vandermonde1");
}

void nmod_poly_mat_mul_vandermonde2(nmod_poly_mat_t res, const nmod_poly_mat_t A,
                                    const nmod_poly_mat_t B)
{
    mul_impl(res, A, B, "nmod_poly_mat_mul_vandermonde2");
}

void nmod_poly_mat_mul_waksman(nmod_poly_mat_t res, const nmod_poly_mat_t A,
                               const nmod_poly_mat_t B)
{
    mul_impl(res, A, B, "nmod_poly_mat_mul_waksman");
}

/* res = (A*B mod x^nhi) div x^nlo, from the first len1 / len2 coefficients of
 * A / B (mulmid.c:28-42 and 470-497: the naive form multiplies in full, so
 * entries are expected to be no longer than len1 / len2). */
static void mulmid_impl(nmod_poly_mat_struct *res, const nmod_poly_mat_struct *A, slong len1,
                        const nmod_poly_mat_struct *B, slong len2, slong nlo, slong nhi,
                        const char *fn)
{
    need_dims(A->c == B->r && res->r == A->r && res->c == B->c, fn);
    if (nlo < 0 || nhi < 0)
        die(fn, "negative coefficient bound");
    if (len1 == 0 || len2 == 0 || nlo >= nhi || A->r == 0 || B->c == 0) {
        pmlg_nmod_poly_mat_zero(res);
        return;
    }
    product_entries(res, A, len1, B, len2, nlo, nhi, fn);
}

void _nmod_poly_mat_mulmid(nmod_poly_mat_t res, const nmod_poly_mat_t A, slong len1,
                           const nmod_poly_mat_t B, slong len2, slong nlo, slong nhi)
{
    const slong l1 = pmlg_nmod_poly_mat_max_length(A), l2 = pmlg_nmod_poly_mat_max_length(B);
    mulmid_impl(res, A, len1 < l1 ? len1 : l1, B, len2 < l2 ? len2 : l2, nlo, nhi,
                "_nmod_poly_mat_mulmid");
}

/* mulmid.c:499-559: nhi is clipped to the product length; the reference's
 * truncations and shifts of the operands only save work, the device middle
 * product sizes its transform from (lenA, lenB, nlo, nhi) directly. */
void nmod_poly_mat_mulmid(nmod_poly_mat_t res, const nmod_poly_mat_t A, const nmod_poly_mat_t B,
                          slong nlo, slong nhi)
{
    const slong len1 = pmlg_nmod_poly_mat_max_length(A), len2 = pmlg_nmod_poly_mat_max_length(B);
    if (len1 > 0 && len2 > 0 && nhi > len1 + len2 - 1)
        nhi = len1 + len2 - 1;
    /* coefficients of degree >= nhi cannot reach the output */
    mulmid_impl(res, A, len1 < nhi ? len1 : nhi, B, len2 < nhi ? len2 : nhi, nlo, nhi,
                "nmod_poly_mat_mulmid");
}

/* ------------------------------------------------------------------------ */
/* approximant bases                                                          */
/* ------------------------------------------------------------------------ */

static void basis_impl(nmod_poly_mat_struct *appbas, slong *shift, const nmod_poly_mat_struct *F,
                       slong order, int iterative, const char *fn)
{
    const slong m = F->r, n = F->c;
    need_dims(appbas->r == m && appbas->c == m, fn);
    if (order < 0)
        die(fn, "negative order");
    /* only coefficients below the order matter (approximant_basis.c:41-42) */
    slong lenF = pmlg_nmod_poly_mat_max_length(F);
    lenF = lenF < order ? lenF : order;
    uint64_t *f = pack(F, lenF, fn);
    const slong cap = order + 1;
    uint64_t *P = xalloc((size_t)(m * m * cap), fn);
    int64_t *sh = (int64_t *)malloc((size_t)(m > 0 ? m : 1) * sizeof(int64_t));
    if (!sh)
        die(fn, "out of host memory");
    for (slong i = 0; i < m; i++)
        sh[i] = shift[i];
    int64_t lenP = 0;
    pmlg_ctx ctx = ctx_acquire(F->modulus, fn);
    if (iterative)
        check(pmlg_mbasis(ctx, P, &lenP, sh, f, m, n, lenF, order, cap), fn);
    else
        check(pmlg_pmbasis(ctx, P, &lenP, sh, f, m, n, lenF, order, cap, PMLG_METHOD_AUTO), fn);
    ctx_release(ctx);
    free(f);
    for (slong i = 0; i < m; i++)
        shift[i] = (slong)sh[i];
    free(sh);
    unpack(appbas, P, cap);
    free(P);
}

void nmod_poly_mat_pmbasis(nmod_poly_mat_t appbas, slong *shift, const nmod_poly_mat_t F,
                           slong order)
{
    basis_impl(appbas, shift, F, order, 0, "nmod_poly_mat_pmbasis");
}

void nmod_poly_mat_mbasis(nmod_poly_mat_t appbas, slong *shift, const nmod_poly_mat_t F,
                          ulong order)
{
    basis_impl(appbas, shift, F, (slong)order, 1, "nmod_poly_mat_mbasis");
}

/* ------------------------------------------------------------------------ */
/* certificate (verification.c:21-116)                                      */
/* ------------------------------------------------------------------------ */

static uint64_t mulmod_h(uint64_t a, uint64_t b, uint64_t p)
{
    return (uint64_t)(((unsigned __int128)a * b) % p);
}

static uint64_t powmod_h(uint64_t a, uint64_t e, uint64_t p)
{
    uint64_t r = 1 % p;
    for (a %= p; e; e >>= 1, a = mulmod_h(a, a, p))
        if (e & 1)
            r = mulmod_h(r, a, p);
    return r;
}

static slong entry_degree(const nmod_poly_struct *e)
{
    slong l = e->length;
    while (l > 0 && e->coeffs[l - 1] == 0)
        l--;
    return l - 1;
}

/* Exact degree of det(P) for a square P (-1 when det = 0): det(P) has degree
 * at most the sum Dmax of the row degrees, so its values at x = 0, 1, ...,
 * Dmax (device evaluations and device determinants) fix it; the degree is
 * the index of the last nonzero Newton divided difference (the Newton basis
 * polynomial of index j has degree j and leading coefficient 1).  Returns -2
 * when the points are not distinct mod p (p <= Dmax) or Dmax >= 4096. */
static slong det_degree(const nmod_poly_mat_struct *P, pmlg_ctx ctx, const char *fn)
{
    const slong m = P->r;
    const uint64_t p = P->modulus;
    slong dmax = 0;
    for (slong i = 0; i < m; i++) {
        slong r = -1;
        for (slong j = 0; j < m; j++) {
            const slong d = entry_degree(nmod_poly_mat_entry(P, i, j));
            r = d > r ? d : r;
        }
        if (r < 0)
            return -1; /* a zero row */
        dmax += r;
    }
    if ((uint64_t)dmax >= p || dmax >= 4096)
        return -2;
    const slong len = pmlg_nmod_poly_mat_max_length(P);
    uint64_t *B = pack(P, len, fn);
    uint64_t *E = xalloc((size_t)(m * m), fn);
    uint64_t *a = xalloc((size_t)(dmax + 1), fn);
    for (slong x = 0; x <= dmax; x++) {
        int64_t rk = 0;
        check(pmlg_polmat_eval(ctx, E, B, m, m, len, (uint64_t)x), fn);
        check(pmlg_const_rank_det(ctx, E, m, m, &rk, &a[x]), fn);
    }
    /* divided differences at the points 0..dmax: x_i - x_{i-j} = j */
    for (slong j = 1; j <= dmax; j++) {
        const uint64_t ij = powmod_h((uint64_t)j, p - 2, p);
        for (slong i = dmax; i >= j; i--)
            a[i] = mulmod_h(a[i] >= a[i - 1] ? a[i] - a[i - 1] : a[i] + (p - a[i - 1]), ij, p);
    }
    slong deg = -1;
    for (slong j = 0; j <= dmax; j++)
        if (a[j])
            deg = j;
    free(a);
    free(E);
    free(B);
    return deg;
}

/* Returns 1 when appbas is a shift-ordered weak Popov approximant basis of F
 * at `order` (ROW_LOWER), else 0, printing which test failed, as the
 * reference does.  The residual appbas*F mod x^(order+1) is a device middle
 * product; the determinant's degree is tested as the reference tests it (see
 * below); the rank of [appbas(0) | coefficient `order` of appbas*F] is a
 * device elimination. */
int nmod_poly_mat_is_approximant_basis(const nmod_poly_mat_t appbas, const nmod_poly_mat_t F,
                                       slong order, const slong *shift, orientation_t orient)
{
    static const char *fn = "nmod_poly_mat_is_approximant_basis";
    const slong m = F->r, n = F->c;
    const uint64_t p = F->modulus;
    int success = 1;
    if (orient != ROW_LOWER)
        die(fn, "only the ROW_LOWER orientation is supported");
    if (appbas->r != m || appbas->c != m) {
        printf("basis has wrong row dimension or column dimension\n");
        return 0;
    }

    /* shifted ordered weak Popov: pivot of row i (rightmost entry reaching the
     * shifted row degree) strictly increasing */
    slong prev = -1;
    for (slong i = 0; i < m; i++) {
        slong piv = -1, mx = 0;
        for (slong j = 0; j < m; j++) {
            const slong d = entry_degree(nmod_poly_mat_entry(appbas, i, j));
            if (d >= 0 && (piv == -1 || mx <= d + shift[j])) {
                piv = j;
                mx = d + shift[j];
            }
        }
        if (piv == -1 || piv <= prev) {
            printf("basis is not shifted-weak Popov\n");
            success = 0;
            break;
        }
        prev = piv;
    }
    const int popov_ok = success;
    if (m == 0)
        return success;

    pmlg_ctx ctx = ctx_acquire(p, fn);
    slong lenP = pmlg_nmod_poly_mat_max_length(appbas);
    slong lenF = pmlg_nmod_poly_mat_max_length(F);
    const slong lenR = order + 1;
    uint64_t *R = xalloc((size_t)(m * n * lenR), fn);
    if (lenP == 0 || lenF == 0 || n == 0) {
        memset(R, 0, (size_t)(m * n * lenR) * sizeof(uint64_t));
    } else {
        lenF = lenF < lenR ? lenF : lenR;
        const slong lp = lenP < lenR ? lenP : lenR;
        uint64_t *a = pack(appbas, lp, fn);
        uint64_t *f = pack(F, lenF, fn);
        check(pmlg_polmat_mulmid(ctx, R, a, f, m, m, n, lp, lenF, 0, lenR, PMLG_METHOD_AUTO), fn);
        free(a);
        free(f);
    }
    for (slong e = 0; e < m * n && success >= 0; e++)
        for (slong d = 0; d < order; d++)
            if (R[e * lenR + d]) {
                printf("entry %ld, %ld of residual has valuation less than the order\n",
                       (long)(e / (n ? n : 1)), (long)(e % (n ? n : 1)));
                success = 0;
                e = m * n;
                break;
            }

    /* deg det(appbas) == D, D the sum of the diagonal degrees
     * (verification.c:73-84 tests the degree only).  A square shifted ordered
     * weak Popov matrix has row i's pivot on the diagonal, so its shifted
     * leading matrix is lower triangular with a nonzero diagonal: it is
     * shift-reduced and deg det = sum_i (deg P_ii + s_i) - sum_i s_i = D.  The
     * determinant is therefore computed only when the weak Popov test failed
     * (the result is already 0 then; only this message depends on it). */
    slong D = 0;
    for (slong i = 0; i < m; i++)
        D += entry_degree(nmod_poly_mat_entry(appbas, i, i));
    if (!popov_ok) {
        const slong dd = det_degree(appbas, ctx, fn);
        if (dd != -2 && dd != D) {
            printf("determinant is not lambda * x**(sum(diag-deg))\n");
            success = 0;
        }
    }

    /* [appbas(0) | C] of full rank, C = coefficient `order` of the residual */
    {
        uint64_t *CP = xalloc((size_t)(m * (m + n)), fn);
        for (slong i = 0; i < m; i++) {
            for (slong j = 0; j < m; j++) {
                const nmod_poly_struct *e = nmod_poly_mat_entry(appbas, i, j);
                CP[i * (m + n) + j] = e->length > 0 ? e->coeffs[0] : 0;
            }
            for (slong j = 0; j < n; j++)
                CP[i * (m + n) + m + j] = R[(i * n + j) * lenR + order];
        }
        int64_t rk = 0;
        check(pmlg_const_rank_det(ctx, CP, m, m + n, &rk, NULL), fn);
        if (rk != m) {
            printf("generation test (step: [C P(0)] has full rank) failed\n");
            success = 0;
        }
        free(CP);
    }
    free(R);
    ctx_release(ctx);
    return success;
}

/* ------------------------------------------------------------------------ */
/* kernel bases (kernel.c:26-478, verification.c:119-215)                    */
/* ------------------------------------------------------------------------ */

/* packed copy of F (r x c) with rows and columns exchanged (transpose) and/or
 * the row order reversed: entry (i, j) of the result is F(ri, cj) where
 * (ri, cj) = (j, i) when transposed, and the result's row i is taken from
 * row dim-1-i when mirrored */
static uint64_t *pack_oriented(const nmod_poly_mat_struct *F, slong len, int transpose, int mirror,
                               const char *fn)
{
    const slong r = transpose ? F->c : F->r, c = transpose ? F->r : F->c;
    uint64_t *b = xalloc((size_t)(r * c * len), fn);
    for (slong i = 0; i < r; i++)
        for (slong j = 0; j < c; j++) {
            const slong si = mirror ? r - 1 - i : i;
            const nmod_poly_struct *e =
                transpose ? nmod_poly_mat_entry(F, j, si) : nmod_poly_mat_entry(F, si, j);
            uint64_t *d = b + (i * c + j) * len;
            const slong l = e->length < len ? e->length : len;
            if (l > 0)
                memcpy(d, e->coeffs, (size_t)l * sizeof(uint64_t));
            if (l < len)
                memset(d + (l > 0 ? l : 0), 0, (size_t)(len - (l > 0 ? l : 0)) * sizeof(uint64_t));
        }
    return b;
}

static void set_entry(nmod_poly_struct *e, const uint64_t *s, slong len)
{
    pthread_once(&g_fit_once, resolve_fit);
    slong l = len;
    while (l > 0 && s[l - 1] == 0)
        l--;
    if (l > 0) {
        g_fit(e, l);
        memcpy(e->coeffs, s, (size_t)l * sizeof(uint64_t));
    }
    e->length = l;
}

/* Left kernel of the packed dim x other matrix g on the device; the basis
 * rows (nul x dim, capacity *cap) are returned in a malloc'd buffer. */
static uint64_t *kernel_device(const uint64_t *g, slong dim, slong other, slong lenF, uint64_t p,
                               int64_t *sh, int64_t *piv, int algo, int64_t *nul, int64_t *cap,
                               const char *fn)
{
    pmlg_ctx ctx = ctx_acquire(p, fn);
    int64_t amp = 0, mn = dim ? sh[0] : 0, mx = mn;
    for (slong i = 0; i < dim; i++) {
        mn = sh[i] < mn ? sh[i] : mn;
        mx = sh[i] > mx ? sh[i] : mx;
    }
    amp = mx - mn;
    int64_t c = (int64_t)((dim < other ? dim : other) + 1) * (lenF > 0 ? lenF : 1) + amp + 1;
    int64_t *sh0 = (int64_t *)malloc((size_t)(dim > 0 ? dim : 1) * sizeof(int64_t));
    if (!sh0)
        die(fn, "out of host memory");
    memcpy(sh0, sh, (size_t)dim * sizeof(int64_t));
    for (;;) {
        uint64_t *K = xalloc((size_t)(dim * dim * c), fn);
        int64_t L = 0;
        memcpy(sh, sh0, (size_t)dim * sizeof(int64_t));
        pmlg_status st = pmlg_kernel_basis(ctx, K, c, &L, nul, piv, sh, g, dim, other, lenF, algo,
                                           PMLG_METHOD_AUTO);
        if (st != PMLG_OK && L > c) {
            free(K);
            c = L;
            continue;
        }
        check(st, fn);
        free(sh0);
        ctx_release(ctx);
        *cap = c;
        return K;
    }
}

/* kernel.c:26-146 for the four orientations, with the ZLS algorithm */
static slong kernel_impl(nmod_poly_mat_struct *ker, slong *pivind, slong *shift,
                         const nmod_poly_mat_struct *pmat, orientation_t orient, int algo,
                         const char *fn)
{
    const int left = orient == ROW_LOWER || orient == ROW_UPPER;
    const int transpose = !left;
    const int mirror = orient == ROW_UPPER || orient == COL_LOWER;
    const slong dim = left ? pmat->r : pmat->c, other = left ? pmat->c : pmat->r;
    if (left)
        need_dims(ker->c == dim && ker->r <= dim, fn);
    else
        need_dims(ker->r == dim && ker->c <= dim, fn);
    const slong lenF = pmlg_nmod_poly_mat_max_length(pmat);
    int64_t *sh = (int64_t *)malloc((size_t)(dim > 0 ? dim : 1) * sizeof(int64_t));
    int64_t *pv = (int64_t *)malloc((size_t)(dim > 0 ? dim : 1) * sizeof(int64_t));
    if (!sh || !pv)
        die(fn, "out of host memory");
    /* the shift, mirrored with the rows; by default the row (left) or column
     * (right) degrees of pmat with zero lines taken as 0 */
    for (slong i = 0; i < dim; i++) {
        const slong si = mirror ? dim - 1 - i : i;
        if (shift) {
            sh[i] = shift[si];
        } else {
            slong d = -1;
            for (slong j = 0; j < other; j++) {
                const slong e = entry_degree(left ? nmod_poly_mat_entry(pmat, si, j)
                                                  : nmod_poly_mat_entry(pmat, j, si));
                d = e > d ? e : d;
            }
            sh[i] = d >= 0 ? d : 0;
        }
    }
    if (dim == 0 || other == 0) {
        /* empty matrix: the identity is a kernel basis (kernel.c:241-249) */
        for (slong i = 0; i < dim; i++)
            for (slong j = 0; j < dim; j++) {
                nmod_poly_struct *e = left ? nmod_poly_mat_entry(ker, i, j)
                                           : nmod_poly_mat_entry(ker, j, i);
                const uint64_t one = 1;
                set_entry(e, &one, i == j ? 1 : 0);
            }
        for (slong i = 0; i < dim; i++)
            if (pivind)
                pivind[i] = i;
        free(sh);
        free(pv);
        return dim;
    }
    uint64_t *g = pack_oriented(pmat, lenF, transpose, mirror, fn);
    int64_t nul = 0, cap = 0;
    uint64_t *K = kernel_device(g, dim, other, lenF, pmat->modulus, sh, pv, algo, &nul, &cap, fn);
    free(g);
    if (left)
        need_dims(ker->r >= nul, fn);
    else
        need_dims(ker->c >= nul, fn);
    /* row t of the device basis becomes row (left) or column (right) t of
     * ker; mirrored orientations reverse the rows, the columns, the shifts
     * and the pivot indices back (kernel.c:82-105, 108-140) */
    for (int64_t t = 0; t < nul; t++) {
        const int64_t st = mirror ? nul - 1 - t : t;
        for (slong j = 0; j < dim; j++) {
            const slong sj = mirror ? dim - 1 - j : j;
            nmod_poly_struct *e = left ? nmod_poly_mat_entry(ker, t, j) : nmod_poly_mat_entry(ker, j, t);
            set_entry(e, K + (st * dim + sj) * cap, cap);
        }
        if (pivind)
            pivind[t] = mirror ? dim - 1 - pv[st] : pv[st];
        if (shift)
            shift[t] = sh[st];
    }
    free(K);
    free(sh);
    free(pv);
    return nul;
}

slong nmod_poly_mat_kernel(nmod_poly_mat_t ker, slong *pivind, slong *shift,
                           const nmod_poly_mat_t pmat, poly_mat_form_t form, orientation_t orient)
{
    static const char *fn = "nmod_poly_mat_kernel";
    if (form > ORD_WEAK_POPOV)
        die(fn, "form > ORD_WEAK_POPOV not implemented.");
    return kernel_impl(ker, pivind, shift, pmat, orient, PMLG_KERNEL_ZLS, fn);
}

slong nmod_poly_mat_kernel_via_approx(nmod_poly_mat_t ker, slong *pivind, slong *shift,
                                      const nmod_poly_mat_t pmat)
{
    return kernel_impl(ker, pivind, shift, pmat, ROW_LOWER, PMLG_KERNEL_VIA_APPROX,
                       "nmod_poly_mat_kernel_via_approx");
}

slong nmod_poly_mat_kernel_zls_approx(nmod_poly_mat_t ker, slong *pivind, slong *shift,
                                      nmod_poly_mat_t pmat)
{
    return kernel_impl(ker, pivind, shift, pmat, ROW_LOWER, PMLG_KERNEL_ZLS,
                       "nmod_poly_mat_kernel_zls_approx");
}

/* pivot index of row (left) or column (right) t of M, as
 * _nmod_poly_vec_pivot_profile with a shift (degrees_pivots_leadingmatrix.c:
 * 276-365): ROW_LOWER/COL_UPPER take the last entry reaching the shifted
 * degree, ROW_UPPER/COL_LOWER the first; -1 for a zero vector */
static slong pivot_of(const nmod_poly_mat_struct *M, slong t, const slong *shift,
                      orientation_t orient)
{
    const int rows = orient == ROW_LOWER || orient == ROW_UPPER;
    const int last = orient == ROW_LOWER || orient == COL_UPPER;
    const slong len = rows ? M->c : M->r;
    slong piv = -1, mx = 0;
    for (slong j = 0; j < len; j++) {
        const slong d = entry_degree(rows ? nmod_poly_mat_entry(M, t, j) : nmod_poly_mat_entry(M, j, t));
        const slong s = shift ? shift[j] : 0;
        if (d >= 0 && (piv == -1 || (last ? mx <= d + s : mx < d + s))) {
            piv = j;
            mx = d + s;
        }
    }
    return piv;
}

/* nmod_poly_mat_is_form.c: weak Popov (distinct pivots, no zero vector) and
 * ordered weak Popov (increasing pivots) */
static int has_form(const nmod_poly_mat_struct *M, const slong *shift, poly_mat_form_t form,
                    orientation_t orient, const char *fn)
{
    if (form == NONE)
        return 1;
    if (form != WEAK_POPOV && form != ORD_WEAK_POPOV)
        die(fn, "only the forms NONE, WEAK_POPOV and ORD_WEAK_POPOV are supported");
    const slong dim = (orient == ROW_LOWER || orient == ROW_UPPER) ? M->r : M->c;
    slong *piv = (slong *)malloc((size_t)(dim > 0 ? dim : 1) * sizeof(slong));
    if (!piv)
        die(fn, "out of host memory");
    int ok = 1;
    for (slong t = 0; t < dim && ok; t++) {
        piv[t] = pivot_of(M, t, shift, orient);
        if (piv[t] < 0)
            ok = 0;
    }
    for (slong t = 0; t < dim && ok; t++)
        for (slong u = (form == ORD_WEAK_POPOV ? t + 1 : 0); u < dim; u++)
            if (u != t && (form == ORD_WEAK_POPOV ? piv[t] >= piv[u] : piv[t] == piv[u])) {
                ok = 0;
                break;
            }
    free(piv);
    return ok;
}

/* Polynomials over Z/pZ on the host for the fraction-free rank below: a
 * coefficient array and its length (no trailing zero coefficient). */
typedef struct {
    uint64_t *c;
    slong len;
} hpoly;

static void hpoly_trim(hpoly *a)
{
    while (a->len > 0 && a->c[a->len - 1] == 0)
        a->len--;
}

/* a*b - c*d over Z/pZ (schoolbook products) */
static hpoly hpoly_cross(const hpoly *a, const hpoly *b, const hpoly *c, const hpoly *d, uint64_t p,
                         const char *fn)
{
    const slong l1 = (a->len && b->len) ? a->len + b->len - 1 : 0;
    const slong l2 = (c->len && d->len) ? c->len + d->len - 1 : 0;
    hpoly r;
    r.len = l1 > l2 ? l1 : l2;
    r.c = xalloc((size_t)(r.len > 0 ? r.len : 1), fn);
    memset(r.c, 0, (size_t)(r.len > 0 ? r.len : 1) * sizeof(uint64_t));
    for (slong i = 0; i < a->len; i++)
        for (slong j = 0; j < b->len; j++) {
            const uint64_t t = mulmod_h(a->c[i], b->c[j], p);
            r.c[i + j] = r.c[i + j] >= p - t ? r.c[i + j] - (p - t) : r.c[i + j] + t;
        }
    for (slong i = 0; i < c->len; i++)
        for (slong j = 0; j < d->len; j++) {
            const uint64_t t = mulmod_h(c->c[i], d->c[j], p);
            r.c[i + j] = r.c[i + j] >= t ? r.c[i + j] - t : r.c[i + j] + (p - t);
        }
    hpoly_trim(&r);
    return r;
}

/* a / b for b nonzero dividing a exactly (schoolbook long division) */
static void hpoly_divexact(hpoly *a, const hpoly *b, uint64_t p, const char *fn)
{
    if (a->len == 0)
        return;
    if (a->len < b->len)
        die(fn, "fraction-free rank: inexact division");
    const uint64_t linv = powmod_h(b->c[b->len - 1], p - 2, p);
    const slong ql = a->len - b->len + 1;
    uint64_t *q = xalloc((size_t)ql, fn);
    for (slong i = ql - 1; i >= 0; i--) {
        const uint64_t f = mulmod_h(a->c[i + b->len - 1], linv, p);
        q[i] = f;
        for (slong j = 0; j < b->len; j++) {
            const uint64_t t = mulmod_h(f, b->c[j], p);
            uint64_t *x = &a->c[i + j];
            *x = *x >= t ? *x - t : *x + (p - t);
        }
    }
    for (slong i = 0; i < b->len - 1; i++)
        if (a->c[i])
            die(fn, "fraction-free rank: inexact division");
    free(a->c);
    a->c = q;
    a->len = ql;
    hpoly_trim(a);
}

/* Rank over Z/p(x) by fraction-free (Bareiss) elimination over Z/pZ[x]: after
 * k pivots every remaining entry is a (k+1)-minor, divisible by the previous
 * pivot, whatever columns were skipped.  Used where Z/pZ has too few points
 * for the rank by evaluation (p <= min(r, c) deg M). */
static slong polmat_rank_bareiss(const nmod_poly_mat_struct *M, const char *fn)
{
    const slong r = M->r, c = M->c;
    const uint64_t p = M->modulus;
    hpoly *A = (hpoly *)malloc((size_t)(r * c > 0 ? r * c : 1) * sizeof(hpoly));
    if (!A)
        die(fn, "out of host memory");
    for (slong i = 0; i < r; i++)
        for (slong j = 0; j < c; j++) {
            const nmod_poly_struct *e = nmod_poly_mat_entry(M, i, j);
            hpoly *h = &A[i * c + j];
            h->len = e->length;
            h->c = xalloc((size_t)(h->len > 0 ? h->len : 1), fn);
            for (slong t = 0; t < h->len; t++)
                h->c[t] = e->coeffs[t] % p;
            hpoly_trim(h);
        }
    uint64_t one = 1;
    hpoly prev = {&one, 1};
    slong rank = 0;
    for (slong col = 0; col < c && rank < r; col++) {
        slong piv = -1;
        for (slong i = rank; i < r && piv < 0; i++)
            if (A[i * c + col].len)
                piv = i;
        if (piv < 0)
            continue;
        for (slong j = 0; j < c; j++) {
            const hpoly t = A[piv * c + j];
            A[piv * c + j] = A[rank * c + j];
            A[rank * c + j] = t;
        }
        const hpoly *pv = &A[rank * c + col];
        for (slong i = rank + 1; i < r; i++) {
            for (slong j = col + 1; j < c; j++) {
                hpoly x = hpoly_cross(pv, &A[i * c + j], &A[i * c + col], &A[rank * c + j], p, fn);
                hpoly_divexact(&x, &prev, p, fn);
                free(A[i * c + j].c);
                A[i * c + j] = x;
            }
            A[i * c + col].len = 0;
        }
        prev = *pv;
        rank++;
    }
    for (slong i = 0; i < r * c; i++)
        free(A[i].c);
    free(A);
    return rank;
}

/* Rank over Z/p(x): a nonzero r x r minor has degree at most r*deg(M), so
 * it does not vanish at one of the points 0, 1, ..., r*deg(M); each point is
 * a device evaluation and a device elimination.  Moduli with fewer points
 * take the fraction-free elimination over Z/pZ[x] on the host. */
static slong polmat_rank(const nmod_poly_mat_struct *M, const char *fn)
{
    const slong r = M->r, c = M->c;
    slong d = -1;
    for (slong i = 0; i < r; i++)
        for (slong j = 0; j < c; j++) {
            const slong e = entry_degree(nmod_poly_mat_entry(M, i, j));
            d = e > d ? e : d;
        }
    if (d < 0)
        return 0;
    const slong full = r < c ? r : c;
    const uint64_t p = M->modulus;
    if ((uint64_t)(full * d) >= p)
        return polmat_rank_bareiss(M, fn);
    pmlg_ctx ctx = ctx_acquire(p, fn);
    const slong len = d + 1;
    uint64_t *P = pack(M, len, fn);
    uint64_t *E = xalloc((size_t)(r * c), fn);
    slong best = 0;
    for (slong x = 0; x <= full * d && best < full; x++) {
        int64_t rk = 0;
        check(pmlg_polmat_eval(ctx, E, P, r, c, len, (uint64_t)x), fn);
        check(pmlg_const_rank_det(ctx, E, r, c, &rk, NULL), fn);
        best = rk > best ? rk : best;
    }
    free(E);
    free(P);
    ctx_release(ctx);
    return best;
}

/* verification.c:119-215: dimensions, nullity = dim - rank, the requested
 * form, and ker * mat == 0 (left) or mat * ker == 0 (right), the product on
 * the device.  As in the reference, generation is not checked. */
int nmod_poly_mat_is_kernel(const nmod_poly_mat_t ker, const slong *shift,
                            const nmod_poly_mat_t mat, poly_mat_form_t form, orientation_t orient)
{
    static const char *fn = "nmod_poly_mat_is_kernel";
    const int left = orient == ROW_LOWER || orient == ROW_UPPER;
    const slong rdim = mat->r, cdim = mat->c;
    int success = 1;
    if (left ? ker->c != rdim : ker->r != cdim) {
        printf(left ? "basis has wrong column dimension\n" : "basis has wrong row dimension\n");
        success = 0;
    }
    const slong rk = polmat_rank(mat, fn);
    if ((left ? ker->r != rdim - rk : ker->c != cdim - rk)) {
        printf("number of rows does not equal nullity\n");
        success = 0;
    }
    if (!has_form(ker, shift, form, orient, fn)) {
        printf("basis does not have the required form\n");
        success = 0;
    }
    if (success && (left ? ker->r : ker->c) > 0) {
        nmod_poly_mat_t res;
        pmlg_nmod_poly_mat_init(res, left ? ker->r : rdim, left ? cdim : ker->c, mat->modulus);
        if (left)
            mul_impl(res, ker, mat, fn);
        else
            mul_impl(res, mat, ker, fn);
        if (pmlg_nmod_poly_mat_max_length(res) != 0) {
            printf("not all rows are in the kernel\n");
            success = 0;
        }
        pmlg_nmod_poly_mat_clear(res);
    }
    return success;
}
