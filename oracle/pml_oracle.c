/*
 * CPU oracle for polynomial-matrix arithmetic over Z/pZ.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): used by tests/, by the smoke check
 * and by the benchmark's CPU-baseline leg; never linked into the product.
 * See pml_oracle.h for the layout conventions and the reference functions
 * each routine restates.
 */
#include "pml_oracle.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* ====================================================================== */
/* scalar arithmetic                                                       */
/* ====================================================================== */

uint64_t orc_mulmod(uint64_t a, uint64_t b, uint64_t p)
{
    return (uint64_t)(((u128)a * b) % p);
}

static inline uint64_t addmod(uint64_t a, uint64_t b, uint64_t p)
{
    /* a,b in [0,p); safe for p up to 2^64-1 */
    return (a >= p - b) ? a - (p - b) : a + b;
}

static inline uint64_t submod(uint64_t a, uint64_t b, uint64_t p)
{
    return (a >= b) ? a - b : a + (p - b);
}

static inline uint64_t negmod(uint64_t a, uint64_t p)
{
    return a == 0 ? 0 : p - a;
}

uint64_t orc_powmod(uint64_t a, uint64_t e, uint64_t p)
{
    uint64_t r = 1 % p;
    a %= p;
    while (e)
    {
        if (e & 1)
            r = orc_mulmod(r, a, p);
        a = orc_mulmod(a, a, p);
        e >>= 1;
    }
    return r;
}

uint64_t orc_invmod(uint64_t a, uint64_t p)
{
    return orc_powmod(a, p - 2, p);
}

int orc_two_adic_valuation(uint64_t n)
{
    return n ? __builtin_ctzll(n) : 64;
}

int orc_is_prime(uint64_t n)
{
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if (n < 2)
        return 0;
    for (int i = 0; i < 12; i++)
    {
        if (n == small[i])
            return 1;
        if (n % small[i] == 0)
            return 0;
    }
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1))
    {
        d >>= 1;
        s++;
    }
    /* this witness set is deterministic for all n < 2^64 */
    static const uint64_t wit[] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    for (int i = 0; i < 7; i++)
    {
        uint64_t a = wit[i] % n;
        if (a == 0)
            continue;
        uint64_t x = orc_powmod(a, d, n);
        if (x == 1 || x == n - 1)
            continue;
        int comp = 1;
        for (int r = 1; r < s; r++)
        {
            x = orc_mulmod(x, x, n);
            if (x == n - 1)
            {
                comp = 0;
                break;
            }
        }
        if (comp)
            return 0;
    }
    return 1;
}

uint64_t orc_root_of_unity_2k(uint64_t p, int k)
{
    if (p < 2 || k < 0 || k > orc_two_adic_valuation(p - 1))
        return 0;
    if (k == 0)
        return 1;
    for (uint64_t g = 2; g < p; g++)
    {
        uint64_t r = orc_powmod(g, (p - 1) >> k, p);
        if (orc_powmod(r, (uint64_t)1 << (k - 1), p) == p - 1)
            return r;
    }
    return 0;
}

/* ====================================================================== */
/* seeded generator                                                        */
/* ====================================================================== */

uint64_t orc_splitmix64(uint64_t *state)
{
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

void orc_rand_uniform(uint64_t *out, int64_t n, uint64_t p, uint64_t seed)
{
    uint64_t st = seed;
    if (p <= 1)
    {
        memset(out, 0, (size_t)n * sizeof(uint64_t));
        return;
    }
    int bits = 64 - __builtin_clzll(p - 1);
    for (int64_t i = 0; i < n; i++)
    {
        uint64_t x;
        do
            x = orc_splitmix64(&st) >> (64 - bits);
        while (x >= p);
        out[i] = x;
    }
}

/* ====================================================================== */
/* lazy 128-bit accumulation of products of residues                       */
/* ====================================================================== */

typedef struct
{
    u128 acc;
    u128 lim; /* acc may receive one more product while acc <= lim */
    uint64_t p;
} acc128;

static inline void acc_init(acc128 *a, uint64_t p)
{
    a->acc = 0;
    a->p = p;
    a->lim = ~(u128)0 - (u128)(p - 1) * (p - 1);
}

static inline void acc_mac(acc128 *a, uint64_t x, uint64_t y)
{
    if (a->acc > a->lim)
        a->acc %= a->p;
    a->acc += (u128)x * y;
}

static inline uint64_t acc_get(const acc128 *a)
{
    return (uint64_t)(a->acc % a->p);
}

/* ====================================================================== */
/* polynomial matrices                                                     */
/* ====================================================================== */

int64_t orc_polmat_max_length(const uint64_t *A, int64_t r, int64_t c, int64_t len)
{
    int64_t best = 0;
    for (int64_t e = 0; e < r * c; e++)
    {
        int64_t l = len;
        while (l > best && A[e * len + l - 1] == 0)
            l--;
        if (l > best)
            best = l;
    }
    return best;
}

void orc_polmat_mul_school(uint64_t *C, const uint64_t *A, const uint64_t *B,
                           int64_t m, int64_t k, int64_t n,
                           int64_t lenA, int64_t lenB, uint64_t p)
{
    if (lenA <= 0 || lenB <= 0)
        return; /* C has padded length 0 */
    const int64_t lenC = lenA + lenB - 1;
    for (int64_t i = 0; i < m; i++)
        for (int64_t j = 0; j < n; j++)
            for (int64_t d = 0; d < lenC; d++)
            {
                acc128 acc;
                acc_init(&acc, p);
                int64_t tlo = d - (lenB - 1) > 0 ? d - (lenB - 1) : 0;
                int64_t thi = d < lenA - 1 ? d : lenA - 1;
                for (int64_t kk = 0; kk < k; kk++)
                {
                    const uint64_t *a = A + (i * k + kk) * lenA;
                    const uint64_t *b = B + (kk * n + j) * lenB;
                    for (int64_t t = tlo; t <= thi; t++)
                        acc_mac(&acc, a[t], b[d - t]);
                }
                C[(i * n + j) * lenC + d] = acc_get(&acc);
            }
}

/* --- radix-2 NTT: decimation in frequency forward (natural order in,
 * bit-reversed order out), decimation in time inverse (bit-reversed in,
 * natural out), so no permutation pass is needed between them; pointwise
 * products work position by position.  Shoup products by the twiddles
 * (Harvey's lazy butterflies, values in [0, 2p), for p < 2^62; reduced after
 * every butterfly above that). --- */

typedef struct
{
    int64_t N;
    int logn;
    uint64_t p;
    int lazy;     /* 4p < 2^64 */
    /* per-stage tables: the stage with half-size h uses w^(j N/(2h)), j < h,
     * stored contiguously from offset N - 2h (h = N/2 at 0, ..., h = 1 at N-2) */
    uint64_t *w, *wq;   /* powers of the root */
    uint64_t *wi, *wiq; /* powers of its inverse */
    uint64_t ninv, ninvq;
} ntt_plan;

static inline uint64_t shoup_pre(uint64_t w, uint64_t p)
{
    return (uint64_t)(((u128)w << 64) / p);
}

/* x * w mod p in [0, 2p) for any x < 2^64 and w < p < 2^63 */
static inline uint64_t shoup_lazy(uint64_t x, uint64_t w, uint64_t wq, uint64_t p)
{
    uint64_t q = (uint64_t)(((u128)x * wq) >> 64);
    return x * w - q * p;
}

/* x * w mod p, for x < 2^64, w < p < 2^63; exact result in [0,p) */
static inline uint64_t shoup_mul(uint64_t x, uint64_t w, uint64_t wq, uint64_t p)
{
    uint64_t r = shoup_lazy(x, w, wq, p);
    return r >= p ? r - p : r;
}

static int ntt_plan_init(ntt_plan *P, int logn, uint64_t p)
{
    if (p >= ((uint64_t)1 << 63))
        return -1;
    uint64_t root = orc_root_of_unity_2k(p, logn);
    if (root == 0 && logn > 0)
        return -1;
    P->N = (int64_t)1 << logn;
    P->logn = logn;
    P->p = p;
    P->lazy = p < ((uint64_t)1 << 62);
    const int64_t N = P->N, sz = N > 1 ? N : 1;
    P->w = malloc(sz * sizeof(uint64_t));
    P->wq = malloc(sz * sizeof(uint64_t));
    P->wi = malloc(sz * sizeof(uint64_t));
    P->wiq = malloc(sz * sizeof(uint64_t));
    const uint64_t rinv = logn > 0 ? orc_invmod(root, p) : 1;
    /* the stage with half-size h uses the (2h)-th root root^(N/2h) */
    for (int64_t h = N / 2; h >= 1; h >>= 1)
    {
        const uint64_t st = orc_powmod(root, (uint64_t)(N / (2 * h)), p);
        const uint64_t sti = orc_powmod(rinv, (uint64_t)(N / (2 * h)), p);
        uint64_t x = 1 % p, y = 1 % p;
        for (int64_t j = 0; j < h; j++)
        {
            P->w[N - 2 * h + j] = x;
            P->wq[N - 2 * h + j] = shoup_pre(x, p);
            P->wi[N - 2 * h + j] = y;
            P->wiq[N - 2 * h + j] = shoup_pre(y, p);
            x = orc_mulmod(x, st, p);
            y = orc_mulmod(y, sti, p);
        }
    }
    P->ninv = orc_invmod((uint64_t)(N % p), p);
    P->ninvq = shoup_pre(P->ninv, p);
    return 0;
}

static void ntt_plan_clear(ntt_plan *P)
{
    free(P->w);
    free(P->wq);
    free(P->wi);
    free(P->wiq);
}

/* forward DIF: a[] in [0, p) natural order -> evaluations in [0, p) at the
 * N-th roots, position t holding the value at root^bitrev(t) */
static void ntt_forward(uint64_t *a, const ntt_plan *P)
{
    const int64_t N = P->N;
    const uint64_t p = P->p, p2 = 2 * p;
    for (int64_t h = N / 2; h >= 1; h >>= 1)
    {
        const uint64_t *w = P->w + (N - 2 * h), *wq = P->wq + (N - 2 * h);
        for (int64_t i = 0; i < N; i += 2 * h)
        {
            uint64_t *x = a + i, *y = a + i + h;
            if (P->lazy)
                for (int64_t j = 0; j < h; j++)
                {
                    const uint64_t u = x[j], v = y[j]; /* [0, 2p) */
                    const uint64_t s = u + v;
                    x[j] = s >= p2 ? s - p2 : s;
                    y[j] = shoup_lazy(u + p2 - v, w[j], wq[j], p);
                }
            else
                for (int64_t j = 0; j < h; j++)
                {
                    const uint64_t u = x[j], v = y[j];
                    x[j] = addmod(u, v, p);
                    y[j] = shoup_mul(submod(u, v, p), w[j], wq[j], p);
                }
        }
    }
    if (P->lazy)
        for (int64_t i = 0; i < N; i++)
            a[i] = a[i] >= p ? a[i] - p : a[i];
}

/* inverse DIT: bit-reversed evaluations in [0, p) -> N^-1 * transform at
 * the inverse root, natural order, in [0, p): the exact inverse of
 * ntt_forward */
static void ntt_inverse(uint64_t *a, const ntt_plan *P)
{
    const int64_t N = P->N;
    const uint64_t p = P->p, p2 = 2 * p;
    for (int64_t h = 1; h < N; h <<= 1)
    {
        const uint64_t *w = P->wi + (N - 2 * h), *wq = P->wiq + (N - 2 * h);
        for (int64_t i = 0; i < N; i += 2 * h)
        {
            uint64_t *x = a + i, *y = a + i + h;
            if (P->lazy)
                for (int64_t j = 0; j < h; j++)
                {
                    uint64_t u = x[j]; /* [0, 2p) */
                    const uint64_t v = shoup_lazy(y[j], w[j], wq[j], p);
                    const uint64_t s = u + v, d = u + p2 - v; /* [0, 4p) */
                    x[j] = s >= p2 ? s - p2 : s;
                    y[j] = d >= p2 ? d - p2 : d;
                }
            else
                for (int64_t j = 0; j < h; j++)
                {
                    const uint64_t u = x[j], v = shoup_mul(y[j], w[j], wq[j], p);
                    x[j] = addmod(u, v, p);
                    y[j] = submod(u, v, p);
                }
        }
    }
    for (int64_t i = 0; i < N; i++)
        a[i] = shoup_mul(a[i], P->ninv, P->ninvq, p);
}

/* ---------------- threads (OpenMP) ---------------- */
static int g_threads = 0; /* 0: OpenMP's default */

void orc_set_threads(int n)
{
    g_threads = n > 0 ? n : 0;
}

int orc_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* Products of residues folded into 128 bits: (p-1)^2 per term, so `fold`
 * terms fit before a reduction (for p < 2^60 that is >= 256). */
static inline int64_t fold_terms(uint64_t p)
{
    const u128 sq = (u128)(p - 1) * (p - 1);
    if (sq == 0)
        return INT64_MAX;
    const u128 f = (~(u128)0 - (u128)(p - 1)) / sq;
    return f > (u128)INT64_MAX ? INT64_MAX : (int64_t)f;
}

int orc_polmat_mul_ntt(uint64_t *C, const uint64_t *A, const uint64_t *B,
                       int64_t m, int64_t k, int64_t n,
                       int64_t lenA, int64_t lenB, uint64_t p)
{
    if (lenA <= 0 || lenB <= 0)
        return 0;
    const int64_t lenC = lenA + lenB - 1;
    int logn = 2; /* at least 4 points: the pointwise loop takes four at a time */
    while (((int64_t)1 << logn) < lenC)
        logn++;
    ntt_plan P;
    if (ntt_plan_init(&P, logn, p))
        return -1;
    const int64_t N = P.N;
    const int nt = orc_threads();
    uint64_t *Ah = malloc((size_t)(m * k) * N * sizeof(uint64_t) + 8);
    uint64_t *Bh = malloc((size_t)(k * n) * N * sizeof(uint64_t) + 8);
    uint64_t *Ch = malloc((size_t)(m * n) * N * sizeof(uint64_t) + 8);
    /* forward transforms, one entry per task */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t e = 0; e < m * k + k * n; e++)
    {
        const int a = e < m * k;
        uint64_t *d = a ? Ah + e * N : Bh + (e - m * k) * N;
        const uint64_t *src = a ? A + e * lenA : B + (e - m * k) * lenB;
        const int64_t l = a ? lenA : lenB;
        memcpy(d, src, (size_t)l * sizeof(uint64_t));
        memset(d + l, 0, (size_t)(N - l) * sizeof(uint64_t));
        ntt_forward(d, &P);
    }
    /* pointwise m x k by k x n products on blocks of SB points: each task
     * copies its block of every operand entry into contiguous panels
     * (A as [i][kk][s], B as [j][kk][s], about 1 MiB), then forms exact
     * 128-bit sums of four points at a time in registers, folded every `fold`
     * terms */
    int64_t SB = 64;
    while (SB > 4 && (m * k + k * n) * SB * 8 > (1 << 20))
        SB >>= 1;
    if (SB > N)
        SB = N;
    const int64_t nsb = N / SB;
    const int64_t fold = fold_terms(p);
#pragma omp parallel num_threads(nt)
    {
        uint64_t *pa = malloc((size_t)(m * k * SB) * sizeof(uint64_t) + 8);
        uint64_t *pb = malloc((size_t)(k * n * SB) * sizeof(uint64_t) + 8);
#pragma omp for schedule(dynamic, 1)
        for (int64_t blk = 0; blk < nsb; blk++)
        {
            const int64_t s0 = blk * SB;
            for (int64_t e = 0; e < m * k; e++)
                memcpy(pa + e * SB, Ah + e * N + s0, (size_t)SB * sizeof(uint64_t));
            for (int64_t kk = 0; kk < k; kk++)
                for (int64_t j = 0; j < n; j++)
                    memcpy(pb + (j * k + kk) * SB, Bh + (kk * n + j) * N + s0,
                           (size_t)SB * sizeof(uint64_t));
            for (int64_t i = 0; i < m; i++)
                for (int64_t j = 0; j < n; j++)
                {
                    uint64_t *out = Ch + (i * n + j) * N + s0;
                    for (int64_t s = 0; s < SB; s += 4)
                    {
                        const uint64_t *a = pa + i * k * SB + s, *b = pb + j * k * SB + s;
                        u128 c0 = 0, c1 = 0, c2 = 0, c3 = 0;
                        for (int64_t kk = 0; kk < k; kk++, a += SB, b += SB)
                        {
                            c0 += (u128)a[0] * b[0];
                            c1 += (u128)a[1] * b[1];
                            c2 += (u128)a[2] * b[2];
                            c3 += (u128)a[3] * b[3];
                            if ((kk + 1) % fold == 0)
                            {
                                c0 %= p;
                                c1 %= p;
                                c2 %= p;
                                c3 %= p;
                            }
                        }
                        out[s] = (uint64_t)(c0 % p);
                        out[s + 1] = (uint64_t)(c1 % p);
                        out[s + 2] = (uint64_t)(c2 % p);
                        out[s + 3] = (uint64_t)(c3 % p);
                    }
                }
        }
        free(pa);
        free(pb);
    }
    /* inverse transforms, one entry per task */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int64_t e = 0; e < m * n; e++)
    {
        ntt_inverse(Ch + e * N, &P);
        memcpy(C + e * lenC, Ch + e * N, (size_t)lenC * sizeof(uint64_t));
    }
    free(Ah);
    free(Bh);
    free(Ch);
    ntt_plan_clear(&P);
    return 0;
}

void orc_polmat_mul(uint64_t *C, const uint64_t *A, const uint64_t *B,
                    int64_t m, int64_t k, int64_t n,
                    int64_t lenA, int64_t lenB, uint64_t p)
{
    if (lenA <= 0 || lenB <= 0)
        return;
    /* small products: the closed form is cheapest */
    if (lenA * lenB <= 4096 || orc_polmat_mul_ntt(C, A, B, m, k, n, lenA, lenB, p) != 0)
        orc_polmat_mul_school(C, A, B, m, k, n, lenA, lenB, p);
}

void orc_polmat_mulmid(uint64_t *out, const uint64_t *A, const uint64_t *B,
                       int64_t m, int64_t k, int64_t n,
                       int64_t lenA, int64_t lenB, int64_t nlo, int64_t nhi, uint64_t p)
{
    const int64_t lout = nhi - nlo > 0 ? nhi - nlo : 0;
    if (lout == 0)
        return;
    memset(out, 0, (size_t)(m * n * lout) * sizeof(uint64_t));
    if (lenA <= 0 || lenB <= 0)
        return;
    const int64_t lenC = lenA + lenB - 1;
    if (nlo >= lenC)
        return;
    /* only coefficients < nhi of the operands can contribute */
    const int64_t la = lenA < nhi ? lenA : nhi;
    const int64_t lb = lenB < nhi ? lenB : nhi;
    if (la <= 0 || lb <= 0)
        return;
    const int64_t lc = la + lb - 1;
    uint64_t *a = malloc((size_t)(m * k * la) * sizeof(uint64_t));
    uint64_t *b = malloc((size_t)(k * n * lb) * sizeof(uint64_t));
    uint64_t *c = malloc((size_t)(m * n * lc) * sizeof(uint64_t));
    for (int64_t e = 0; e < m * k; e++)
        memcpy(a + e * la, A + e * lenA, la * sizeof(uint64_t));
    for (int64_t e = 0; e < k * n; e++)
        memcpy(b + e * lb, B + e * lenB, lb * sizeof(uint64_t));
    orc_polmat_mul(c, a, b, m, k, n, la, lb, p);
    for (int64_t e = 0; e < m * n; e++)
        for (int64_t d = 0; d < lout; d++)
            if (nlo + d < lc)
                out[e * lout + d] = c[e * lc + nlo + d];
    free(a);
    free(b);
    free(c);
}

void orc_polmat_eval(uint64_t *out, const uint64_t *A, int64_t r, int64_t c, int64_t len,
                     uint64_t x, uint64_t p)
{
    x %= p;
    /* Horner with Shoup products by the fixed point (p < 2^63), else exact
     * 128-bit remainders; one entry per task */
    const int shoup = p < ((uint64_t)1 << 63);
    const uint64_t xq = shoup ? shoup_pre(x, p) : 0;
    const int nt = r * c * len >= (1 << 16) ? orc_threads() : 1;
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int64_t e = 0; e < r * c; e++)
    {
        uint64_t v = 0;
        const uint64_t *a = A + e * len;
        if (shoup)
            for (int64_t d = len - 1; d >= 0; d--)
                v = addmod(shoup_mul(v, x, xq, p), a[d], p);
        else
            for (int64_t d = len - 1; d >= 0; d--)
                v = addmod(orc_mulmod(v, x, p), a[d], p);
        out[e] = v;
    }
}

/* ====================================================================== */
/* constant matrices                                                       */
/* ====================================================================== */

int64_t orc_rref(uint64_t *M, int64_t r, int64_t c, uint64_t p, int64_t *pivcols)
{
    int64_t rank = 0;
    for (int64_t col = 0; col < c && rank < r; col++)
    {
        int64_t piv = -1;
        for (int64_t row = rank; row < r; row++)
            if (M[row * c + col] != 0)
            {
                piv = row;
                break;
            }
        if (piv < 0)
            continue;
        if (piv != rank)
            for (int64_t cc = 0; cc < c; cc++)
            {
                uint64_t t = M[piv * c + cc];
                M[piv * c + cc] = M[rank * c + cc];
                M[rank * c + cc] = t;
            }
        uint64_t inv = orc_invmod(M[rank * c + col], p);
        for (int64_t cc = col; cc < c; cc++)
            M[rank * c + cc] = orc_mulmod(M[rank * c + cc], inv, p);
        for (int64_t row = 0; row < r; row++)
        {
            if (row == rank)
                continue;
            uint64_t f = M[row * c + col];
            if (f == 0)
                continue;
            for (int64_t cc = col; cc < c; cc++)
                M[row * c + cc] = submod(M[row * c + cc], orc_mulmod(f, M[rank * c + cc], p), p);
        }
        if (pivcols)
            pivcols[rank] = col;
        rank++;
    }
    return rank;
}

int64_t orc_rank(const uint64_t *A, int64_t r, int64_t c, uint64_t p)
{
    uint64_t *M = malloc((size_t)(r * c + 1) * sizeof(uint64_t));
    memcpy(M, A, (size_t)(r * c) * sizeof(uint64_t));
    int64_t rk = orc_rref(M, r, c, p, NULL);
    free(M);
    return rk;
}

/* Restates left_nullspace.c:42-102 on top of FLINT's nullspace convention
 * (right kernel of A^T from its reduced row echelon form R: basis vector i
 * has 1 at the i-th non-pivot column and -R[j][nonpivot_i] at pivot j).
 * PML keeps, for each basis vector, the entries at the pivot columns of R
 * (its "row rank profile"), giving X[i][j] = -R[j][nonpivot_i]. */
int64_t orc_left_nullspace_compact(uint64_t *X, int64_t *permutation,
                                   const uint64_t *A, int64_t m, int64_t n, uint64_t p)
{
    /* At = transpose(A), n x m */
    uint64_t *At = malloc((size_t)(n * m + 1) * sizeof(uint64_t));
    int64_t *piv = malloc((size_t)(m + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < m; i++)
        for (int64_t j = 0; j < n; j++)
            At[j * m + i] = A[i * n + j];
    const int64_t rank = orc_rref(At, n, m, p, piv);
    const int64_t nullity = m - rank;
    if (nullity == 0)
    {
        for (int64_t i = 0; i < m; i++)
            permutation[i] = i;
    }
    else
    {
        int64_t np = 0, q = 0;
        for (int64_t i = 0; i < m; i++)
        {
            if (q < rank && piv[q] == i)
            {
                permutation[q] = i;
                q++;
            }
            else
            {
                permutation[rank + np] = i;
                np++;
            }
        }
        for (int64_t i = 0; i < nullity; i++)
            for (int64_t j = 0; j < rank; j++)
                X[i * rank + j] = negmod(At[j * m + permutation[rank + i]], p);
    }
    free(At);
    free(piv);
    return nullity;
}

/* ====================================================================== */
/* M-Basis                                                                 */
/* ====================================================================== */

typedef struct
{
    int64_t value, index;
} pair64;

static int pair_cmp(const void *a, const void *b)
{
    const pair64 *x = a, *y = b;
    if (x->value != y->value)
        return x->value < y->value ? -1 : 1;
    return (x->index > y->index) - (x->index < y->index);
}

/* entry-major (r x c x len) -> coefficient-major (len x r x c), truncated to L */
static void to_coeff_major(uint64_t *dst, const uint64_t *src, int64_t r, int64_t c,
                           int64_t len, int64_t L)
{
    for (int64_t d = 0; d < L; d++)
        for (int64_t e = 0; e < r * c; e++)
            dst[d * r * c + e] = src[e * len + d];
}

void orc_mbasis(uint64_t **Pout, int64_t *lenP, int64_t *shift,
                const uint64_t *F, int64_t m, int64_t n, int64_t lenF,
                int64_t order, uint64_t p)
{
    /* matp = F truncated to `order`, normalised (nmod_mat_poly_set_trunc_from_poly_mat) */
    int64_t lt = orc_polmat_max_length(F, m, n, lenF);
    if (lt > order)
        lt = order;
    if (lt < 0)
        lt = 0;
    uint64_t *Fc = malloc((size_t)(lt * m * n + 1) * sizeof(uint64_t));
    to_coeff_major(Fc, F, m, n, lenF, lt);
    while (lt > 0)
    {
        int z = 1;
        for (int64_t e = 0; e < m * n; e++)
            if (Fc[(lt - 1) * m * n + e])
            {
                z = 0;
                break;
            }
        if (!z)
            break;
        lt--;
    }

    /* appbas = identity; its length grows by at most one per step, or by
     * order-ord once when the residual has full row rank */
    const int64_t cap = order + 2;
    const int64_t mm = m * m;
    uint64_t *app = calloc((size_t)(cap * mm + 1), sizeof(uint64_t));
    int64_t L = 1;
    for (int64_t i = 0; i < m; i++)
        app[i * m + i] = 1 % p;

    if (lt > 0)
    {
        uint64_t *res = malloc((size_t)(m * n + 1) * sizeof(uint64_t));
        uint64_t *resp = malloc((size_t)(m * n + 1) * sizeof(uint64_t));
        uint64_t *ns = malloc((size_t)(m * m + 1) * sizeof(uint64_t));
        uint64_t *row = malloc((size_t)(m + 1) * sizeof(uint64_t));
        uint64_t *tmp = malloc((size_t)(cap * mm + 1) * sizeof(uint64_t));
        int64_t *perm = malloc((size_t)(m + 1) * sizeof(int64_t));
        int64_t *piv = malloc((size_t)(m + 1) * sizeof(int64_t));
        pair64 *pairs = malloc((size_t)(m + 1) * sizeof(pair64));

        for (int64_t ord = 0; ord < order; ord++)
        {
            /* residual = coefficient ord of appbas * matp (nmod_mat_poly_mul_coeff) */
            const int64_t ub = ord + 1 < L ? ord + 1 : L;
            const int64_t lb = ord + 1 - lt > 0 ? ord + 1 - lt : 0;
            for (int64_t i = 0; i < m; i++)
                for (int64_t j = 0; j < n; j++)
                {
                    acc128 acc;
                    acc_init(&acc, p);
                    for (int64_t t = lb; t < ub; t++)
                    {
                        const uint64_t *a = app + t * mm + i * m;
                        const uint64_t *f = Fc + (ord - t) * m * n + j;
                        for (int64_t kk = 0; kk < m; kk++)
                            acc_mac(&acc, a[kk], f[kk * n]);
                    }
                    res[i * n + j] = acc_get(&acc);
                }

            /* stable permutation sorting (shift, index) increasingly */
            for (int64_t i = 0; i < m; i++)
            {
                pairs[i].value = shift[i];
                pairs[i].index = i;
            }
            qsort(pairs, (size_t)m, sizeof(pair64), pair_cmp);
            for (int64_t i = 0; i < m; i++)
                perm[i] = pairs[i].index;
            for (int64_t i = 0; i < m; i++)
                memcpy(resp + i * n, res + perm[i] * n, (size_t)n * sizeof(uint64_t));

            const int64_t nullity = orc_left_nullspace_compact(ns, piv, resp, m, n, p);
            const int64_t rank = m - nullity;

            if (nullity == 0)
            {
                /* appbas <- X^(order-ord) * appbas and stop */
                const int64_t s = order - ord;
                memmove(app + s * mm, app, (size_t)(L * mm) * sizeof(uint64_t));
                memset(app, 0, (size_t)(s * mm) * sizeof(uint64_t));
                L += s;
                for (int64_t i = 0; i < m; i++)
                    shift[i] += s;
                break;
            }
            if (nullity == m)
                continue; /* residual is zero: nothing to update */

            /* composed permutation: pivots[i] <- perm[pivots[i]] */
            for (int64_t i = 0; i < m; i++)
                piv[i] = perm[piv[i]];
            for (int64_t i = 0; i < rank; i++)
                shift[piv[i]] += 1;

            /* permuted world: row i of tmp = row piv[i] of appbas */
            for (int64_t d = 0; d < L; d++)
                for (int64_t i = 0; i < m; i++)
                    memcpy(tmp + d * mm + i * m, app + d * mm + piv[i] * m,
                           (size_t)m * sizeof(uint64_t));

            /* constant update: rows rank.. += nsbas * rows 0..rank-1 */
            for (int64_t d = 0; d < L; d++)
            {
                uint64_t *cf = tmp + d * mm;
                for (int64_t t = 0; t < nullity; t++)
                {
                    for (int64_t j = 0; j < m; j++)
                    {
                        acc128 acc;
                        acc_init(&acc, p);
                        for (int64_t r = 0; r < rank; r++)
                            acc_mac(&acc, ns[t * rank + r], cf[r * m + j]);
                        row[j] = addmod(cf[(rank + t) * m + j], acc_get(&acc), p);
                    }
                    memcpy(cf + (rank + t) * m, row, (size_t)m * sizeof(uint64_t));
                }
            }

            /* length grows if one of the rows to be multiplied by X reaches the top */
            int grow = 0;
            for (int64_t i = 0; i < rank && !grow; i++)
                for (int64_t j = 0; j < m; j++)
                    if (tmp[(L - 1) * mm + i * m + j])
                    {
                        grow = 1;
                        break;
                    }
            if (grow)
            {
                memset(tmp + L * mm, 0, (size_t)mm * sizeof(uint64_t));
                L++;
            }
            /* rows 0..rank-1 multiplied by X */
            for (int64_t d = L - 1; d > 0; d--)
                for (int64_t i = 0; i < rank; i++)
                    memcpy(tmp + d * mm + i * m, tmp + (d - 1) * mm + i * m,
                           (size_t)m * sizeof(uint64_t));
            for (int64_t i = 0; i < rank; i++)
                memset(tmp + i * m, 0, (size_t)m * sizeof(uint64_t));

            /* back from the permuted world: appbas row piv[i] = tmp row i */
            for (int64_t d = 0; d < L; d++)
                for (int64_t i = 0; i < m; i++)
                    memcpy(app + d * mm + piv[i] * m, tmp + d * mm + i * m,
                           (size_t)m * sizeof(uint64_t));
        }
        free(res);
        free(resp);
        free(ns);
        free(row);
        free(tmp);
        free(perm);
        free(piv);
        free(pairs);
    }
    free(Fc);

    /* coefficient-major -> entry-major */
    uint64_t *P = malloc((size_t)(mm * L + 1) * sizeof(uint64_t));
    for (int64_t d = 0; d < L; d++)
        for (int64_t e = 0; e < mm; e++)
            P[e * L + d] = app[d * mm + e];
    free(app);
    *Pout = P;
    *lenP = L;
}

/* ====================================================================== */
/* PM-Basis                                                                */
/* ====================================================================== */

#define ORC_PMBASIS_THRES 32 /* PMBASIS_THRES, nmod_poly_mat_approximant.h:101 */

void orc_pmbasis(uint64_t **Pout, int64_t *lenP, int64_t *shift,
                 const uint64_t *F, int64_t m, int64_t n, int64_t lenF,
                 int64_t order, uint64_t p)
{
    /* pmat = F truncated to order (approximant_basis.c:39-42) */
    int64_t lt = lenF < order ? lenF : order;
    if (lt < 0)
        lt = 0;
    uint64_t *pm = malloc((size_t)(m * n * lt + 1) * sizeof(uint64_t));
    for (int64_t e = 0; e < m * n; e++)
        memcpy(pm + e * lt, F + e * lenF, (size_t)lt * sizeof(uint64_t));

    if (order <= ORC_PMBASIS_THRES)
    {
        orc_mbasis(Pout, lenP, shift, pm, m, n, lt, order, p);
        free(pm);
        return;
    }

    const int64_t order1 = order >> 1;
    const int64_t order2 = order - order1;

    uint64_t *P1;
    int64_t L1;
    orc_pmbasis(&P1, &L1, shift, pm, m, n, lt, order1, p);

    /* residual = (P1 * pmat mod x^order) div x^order1 */
    uint64_t *res = malloc((size_t)(m * n * order2 + 1) * sizeof(uint64_t));
    orc_polmat_mulmid(res, P1, pm, m, m, n, L1, lt, order1, order, p);

    uint64_t *P2;
    int64_t L2;
    orc_pmbasis(&P2, &L2, shift, res, m, n, order2, order2, p);

    /* appbas = P2 * P1 */
    const int64_t L = L1 + L2 - 1;
    uint64_t *P = malloc((size_t)(m * m * L + 1) * sizeof(uint64_t));
    orc_polmat_mul(P, P2, P1, m, m, m, L2, L1, p);

    free(P1);
    free(P2);
    free(res);
    free(pm);
    *Pout = P;
    *lenP = L;
}

/* ====================================================================== */
/* certificate                                                             */
/* ====================================================================== */

static int64_t poly_degree(const uint64_t *a, int64_t len)
{
    int64_t d = len - 1;
    while (d >= 0 && a[d] == 0)
        d--;
    return d;
}

/* determinant of an r x r constant matrix (Gaussian elimination) */
static uint64_t det_const(uint64_t *M, int64_t r, uint64_t p)
{
    uint64_t det = 1 % p;
    for (int64_t c = 0; c < r; c++)
    {
        int64_t piv = -1;
        for (int64_t i = c; i < r; i++)
            if (M[i * r + c])
            {
                piv = i;
                break;
            }
        if (piv < 0)
            return 0;
        if (piv != c)
        {
            for (int64_t j = 0; j < r; j++)
            {
                uint64_t t = M[piv * r + j];
                M[piv * r + j] = M[c * r + j];
                M[c * r + j] = t;
            }
            det = negmod(det, p);
        }
        det = orc_mulmod(det, M[c * r + c], p);
        uint64_t inv = orc_invmod(M[c * r + c], p);
        for (int64_t i = c + 1; i < r; i++)
        {
            uint64_t f = orc_mulmod(M[i * r + c], inv, p);
            if (!f)
                continue;
            for (int64_t j = c; j < r; j++)
                M[i * r + j] = submod(M[i * r + j], orc_mulmod(f, M[c * r + j], p), p);
        }
    }
    return det;
}

int orc_is_approximant_basis(const uint64_t *P, int64_t lenP,
                             const uint64_t *F, int64_t m, int64_t n, int64_t lenF,
                             int64_t order, const int64_t *shift, uint64_t p,
                             uint64_t seed)
{
    int fail = 0;

    /* shifted ordered weak Popov, ROW_LOWER convention
     * (degrees_pivots_leadingmatrix.c: rightmost entry reaching the max) */
    int64_t prev = -1;
    for (int64_t i = 0; i < m; i++)
    {
        int64_t piv = -1, mx = 0;
        for (int64_t j = 0; j < m; j++)
        {
            int64_t d = poly_degree(P + (i * m + j) * lenP, lenP);
            if (d >= 0 && (piv == -1 || mx <= d + shift[j]))
            {
                piv = j;
                mx = d + shift[j];
            }
        }
        if (piv == -1 || piv <= prev)
        {
            fail |= 1;
            break;
        }
        prev = piv;
    }

    /* residual P*F: coefficients [0, order] */
    uint64_t *R = malloc((size_t)(m * n * (order + 1) + 1) * sizeof(uint64_t));
    orc_polmat_mulmid(R, P, F, m, m, n, lenP, lenF, 0, order + 1, p);
    for (int64_t e = 0; e < m * n && !(fail & 2); e++)
        for (int64_t d = 0; d < order; d++)
            if (R[e * (order + 1) + d])
            {
                fail |= 2;
                break;
            }

    /* deg det(P) == D with D the sum of the diagonal degrees, the degree
     * test of verification.c:73-84 (not a test that det is a monomial).
     * det(P) has degree at most Dmax, the sum of the row degrees, so its
     * values at 0..Dmax fix it: the degree is the index of the last nonzero
     * Newton divided difference.  A shifted ordered weak Popov P has
     * deg det = D by construction (its shifted leading matrix is lower
     * triangular with a nonzero diagonal), so the values are computed only
     * for a P that failed that test, when p > Dmax (distinct points) and the
     * work is small; otherwise the degree is left undecided (bit 4 not set). */
    if (m > 0)
    {
        int64_t D = 0, dmax = 0, zero_row = 0;
        for (int64_t i = 0; i < m; i++)
        {
            D += poly_degree(P + (i * m + i) * lenP, lenP);
            int64_t r = -1;
            for (int64_t j = 0; j < m; j++)
            {
                const int64_t d = poly_degree(P + (i * m + j) * lenP, lenP);
                r = d > r ? d : r;
            }
            if (r < 0)
                zero_row = 1;
            dmax += r > 0 ? r : 0;
        }
        int64_t deg = -2;
        if (!(fail & 1))
            deg = D;
        else if (zero_row)
            deg = -1;
        else if ((uint64_t)dmax < p &&
                 (double)(dmax + 1) * (double)(m * m) * (double)(lenP + m) < 0x1p30)
        {
            uint64_t *E = malloc((size_t)(m * m) * sizeof(uint64_t));
            uint64_t *a = malloc((size_t)(dmax + 1) * sizeof(uint64_t));
            for (int64_t x = 0; x <= dmax; x++)
            {
                orc_polmat_eval(E, P, m, m, lenP, (uint64_t)x, p);
                a[x] = det_const(E, m, p);
            }
            for (int64_t j = 1; j <= dmax; j++)
            {
                const uint64_t ij = orc_invmod((uint64_t)j, p);
                for (int64_t i = dmax; i >= j; i--)
                    a[i] = orc_mulmod(submod(a[i], a[i - 1], p), ij, p);
            }
            deg = -1;
            for (int64_t j = 0; j <= dmax; j++)
                if (a[j])
                    deg = j;
            free(a);
            free(E);
        }
        if (deg != -2 && deg != D)
            fail |= 4;
    }

    /* [P(0) | C] full rank, C = coefficient `order` of P*F */
    uint64_t *CP = malloc((size_t)(m * (m + n) + 1) * sizeof(uint64_t));
    for (int64_t i = 0; i < m; i++)
    {
        for (int64_t j = 0; j < m; j++)
            CP[i * (m + n) + j] = lenP > 0 ? P[(i * m + j) * lenP] : 0;
        for (int64_t j = 0; j < n; j++)
            CP[i * (m + n) + m + j] = R[(i * n + j) * (order + 1) + order];
    }
    if (orc_rank(CP, m, m + n, p) != m)
        fail |= 8;
    free(CP);
    free(R);
    return fail;
}

void orc_free(void *ptr)
{
    free(ptr);
}
