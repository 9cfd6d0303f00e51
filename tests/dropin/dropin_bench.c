/*
 * Timing of the FLINT-layout product path at 32x32 by 32x32, length 4096,
 * p = 2^60 - 3*2^24 + 1: what a PML caller of nmod_poly_mat_multiply pays,
 * host memory to host memory.
 *
 *   dropin      nmod_poly_mat_multiply: entries read in place and written in
 *               place through pinned, multi-threaded, pipelined staging
 *               (pmlg_polmat_mul_entries);
 *   packed      the same product through a packed host copy: pack the FLINT
 *               entries into one buffer, pmlg_polmat_mul (pageable copies),
 *               unpack into the result entries;
 *   resident    pmlg_polmat_mul_dev on inputs already in device memory (the
 *               device work alone).
 *
 * Prints one JSON line.  Checks that the drop-in and packed results agree.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "pml_oracle.h"
#include "pmlgpu_flint.h"

static double now(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void fill(nmod_poly_mat_t M, slong len, uint64_t seed)
{
    for (slong i = 0; i < M->r; i++)
        for (slong j = 0; j < M->c; j++) {
            nmod_poly_struct *e = nmod_poly_mat_entry(M, i, j);
            nmod_poly_fit_length(e, len);
            orc_rand_uniform(e->coeffs, len, M->modulus, seed + (uint64_t)(i * M->c + j));
            e->coeffs[len - 1] |= 1;
            e->length = len;
        }
}

static uint64_t *pack(const nmod_poly_mat_t M, slong len)
{
    uint64_t *b = malloc((size_t)(M->r * M->c * len) * 8);
    for (slong i = 0; i < M->r; i++)
        for (slong j = 0; j < M->c; j++) {
            const nmod_poly_struct *e = nmod_poly_mat_entry(M, i, j);
            memcpy(b + (i * M->c + j) * len, e->coeffs, (size_t)e->length * 8);
            memset(b + (i * M->c + j) * len + e->length, 0, (size_t)(len - e->length) * 8);
        }
    return b;
}

static void unpack(nmod_poly_mat_t M, const uint64_t *b, slong len)
{
    for (slong i = 0; i < M->r; i++)
        for (slong j = 0; j < M->c; j++) {
            nmod_poly_struct *e = nmod_poly_mat_entry(M, i, j);
            const uint64_t *s = b + (i * M->c + j) * len;
            slong l = len;
            while (l > 0 && s[l - 1] == 0)
                l--;
            nmod_poly_fit_length(e, l);
            memcpy(e->coeffs, s, (size_t)l * 8);
            e->length = l;
        }
}

int main(int argc, char **argv)
{
    const int reps = argc > 1 ? atoi(argv[1]) : 10;
    const uint64_t p = 1152921504556515329ULL;
    const slong M = 32, L = 4096, LC = 2 * L - 1;
    nmod_poly_mat_t A, B, C, D;
    nmod_poly_mat_init(A, M, M, p);
    nmod_poly_mat_init(B, M, M, p);
    nmod_poly_mat_init(C, M, M, p);
    nmod_poly_mat_init(D, M, M, p);
    fill(A, L, 1);
    fill(B, L, 100000);

    /* drop-in (warm-up creates the context, tables and stagers) */
    nmod_poly_mat_multiply(C, A, B);
    double t0 = now();
    for (int r = 0; r < reps; r++)
        nmod_poly_mat_multiply(C, A, B);
    const double t_drop = (now() - t0) / reps;

    /* packed host copy, pageable transfers */
    pmlg_ctx ctx;
    if (pmlg_ctx_create(&ctx, p, 0) != PMLG_OK) {
        printf("{\"error\": \"%s\"}\n", pmlg_last_error());
        return 1;
    }
    uint64_t *c = malloc((size_t)(M * M * LC) * 8);
    double t_pack = 0;
    for (int r = -1; r < reps; r++) {
        t0 = now();
        uint64_t *a = pack(A, L), *b = pack(B, L);
        if (pmlg_polmat_mul(ctx, c, a, b, M, M, M, L, L, PMLG_METHOD_AUTO) != PMLG_OK) {
            printf("{\"error\": \"%s\"}\n", pmlg_last_error());
            return 1;
        }
        unpack(D, c, LC);
        free(a);
        free(b);
        if (r >= 0)
            t_pack += now() - t0;
    }
    t_pack /= reps;
    const int same = nmod_poly_mat_equal(C, D);

    /* device-resident */
    void *dA, *dB, *dC;
    uint64_t *a = pack(A, L), *b = pack(B, L);
    pmlg_dev_alloc(ctx, &dA, (uint64_t)(M * M * L) * 8);
    pmlg_dev_alloc(ctx, &dB, (uint64_t)(M * M * L) * 8);
    pmlg_dev_alloc(ctx, &dC, (uint64_t)(M * M * LC) * 8);
    pmlg_memcpy_h2d(ctx, dA, a, (uint64_t)(M * M * L) * 8);
    pmlg_memcpy_h2d(ctx, dB, b, (uint64_t)(M * M * L) * 8);
    pmlg_polmat_mul_dev(ctx, dC, dA, dB, M, M, M, L, L, PMLG_METHOD_AUTO, NULL);
    pmlg_ctx_sync(ctx);
    t0 = now();
    for (int r = 0; r < reps; r++)
        pmlg_polmat_mul_dev(ctx, dC, dA, dB, M, M, M, L, L, PMLG_METHOD_AUTO, NULL);
    pmlg_ctx_sync(ctx);
    const double t_res = (now() - t0) / reps;

    const double mb = (double)(2 * M * M * L + M * M * LC) * 8 / 1e6;
    printf("{\"workload\": \"32x32*32x32 len 4096 p=2^60-3*2^24+1, FLINT-layout host matrices\", "
           "\"dropin_ms\": %.3f, \"packed_ms\": %.3f, \"resident_ms\": %.3f, "
           "\"host_bytes_MB\": %.1f, \"dropin_host_GBps\": %.2f, \"packed_host_GBps\": %.2f, "
           "\"results_equal\": %s, \"reps\": %d}\n",
           t_drop * 1e3, t_pack * 1e3, t_res * 1e3, mb, mb / 1e3 / t_drop, mb / 1e3 / t_pack,
           same ? "true" : "false", reps);
    pmlg_dev_free(ctx, dA);
    pmlg_dev_free(ctx, dB);
    pmlg_dev_free(ctx, dC);
    pmlg_ctx_destroy(ctx);
    free(a);
    free(b);
    free(c);
    nmod_poly_mat_clear(A);
    nmod_poly_mat_clear(B);
    nmod_poly_mat_clear(C);
    nmod_poly_mat_clear(D);
    return same ? 0 : 1;
}
