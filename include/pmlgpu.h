/*
 * pmlgpu: polynomial-matrix multiplication over Z/pZ on one MI355X (gfx950).
 *
 * C ABI of the device library libpmlgpu.so.  Two layers:
 *
 *  1. The explicit status API (pmlg_*): packed, entry-major polynomial
 *     matrices (entry (i,j) coefficient d of an r x c matrix of padded length
 *     len at a[(i*c + j)*len + d], residues in [0,p)), with host-pointer and
 *     device-pointer variants.  Every call returns a pmlg_status; the text of
 *     the last failure is available from pmlg_last_error().
 *
 *  2. The FLINT-shaped drop-in (pmlgpu_flint.h): nmod_poly_mat_multiply,
 *     nmod_poly_mat_mulmid, nmod_poly_mat_pmbasis, ... on nmod_poly_mat_t.
 *
 * Reference interfaces replaced (vneiger/pml, flint-extras/src):
 *   pmlg_polmat_mul*     <- nmod_poly_mat_multiply    nmod_poly_mat_multiply.h:92
 *                           (and the evaluation back-ends it dispatches to,
 *                            nmod_poly_mat_mul_geometric :80, _vandermonde1/2 :63-64,
 *                            _waksman :72, whose result is the same unique product)
 *   pmlg_polmat_mulmid*  <- nmod_poly_mat_mulmid      nmod_poly_mat_multiply.h:122-124
 *                           _nmod_poly_mat_mulmid     nmod_poly_mat_multiply.h:128-131
 *   pmlg_pmbasis         <- nmod_poly_mat_pmbasis     nmod_poly_mat_approximant.h:282-285
 *   pmlg_mbasis          <- nmod_poly_mat_mbasis      nmod_poly_mat_approximant.h:321-324
 *   pmlg_kernel_basis    <- nmod_poly_mat_kernel_zls_approx / _via_approx
 *                                                     nmod_poly_mat_kernel.h:91-105
 *   pmlg_polmat_mul_batch: no single reference counterpart; it runs many
 *                          independent nmod_poly_mat_multiply calls at once.
 *
 * Supported moduli: any p in [2, 2^64).  When p < 2^62 is prime and
 * 2^ceil(log2(N)) | p-1 for the transform length N a product needs (N <= 2^24
 * for p >= 2^30, 2^26 for p < 2^30), the product is evaluated modulo p
 * itself; otherwise (non-FFT primes, composite moduli, p >= 2^62) it is
 * computed modulo three 60-bit FFT primes and recombined by CRT (N <= 2^24,
 * k min(lenA, lenB) < 2^51).  Approximant and kernel bases need a prime modulus
 * (p = 2 included).
 * Longer products return PMLG_ERR_UNSUPPORTED.
 */
#ifndef PMLGPU_H
#define PMLGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    PMLG_OK = 0,
    PMLG_ERR_ARG = 1,          /* invalid argument (dimension, length, pointer) */
    PMLG_ERR_UNSUPPORTED = 2,  /* modulus / length / shape outside what the device path handles */
    PMLG_ERR_DEVICE = 3,       /* HIP runtime error */
    PMLG_ERR_ALLOC = 4         /* device or host allocation failed */
} pmlg_status;

/* Pointwise-product kernel selection. */
typedef enum {
    PMLG_METHOD_AUTO = 0,  /* MFMA when the shape and modulus allow it, plain otherwise */
    PMLG_METHOD_PLAIN = 1, /* per-lane exact 128-bit (or 64-bit) dot products */
    PMLG_METHOD_MFMA = 2,  /* int8-limb MFMA contraction with int32 accumulation */
    PMLG_METHOD_THRESHOLD = 3 /* batched products: the hand-set threshold rule only */
} pmlg_method;

typedef struct pmlg_ctx_s *pmlg_ctx;

/* A context holds one prime, its twiddle tables, a HIP stream and scratch
 * buffers on one device.  Calls on one context are serialised internally. */
pmlg_status pmlg_ctx_create(pmlg_ctx *ctx, uint64_t p, int device);
void pmlg_ctx_destroy(pmlg_ctx ctx);
/* log2 of the largest transform length evaluated modulo p itself (-1 when
 * every product takes the multi-prime path, e.g. for a composite modulus) */
int pmlg_ctx_max_log_len(pmlg_ctx ctx);
/* Stream used by the host-pointer calls (a hipStream_t). */
void *pmlg_ctx_stream(pmlg_ctx ctx);
pmlg_status pmlg_ctx_sync(pmlg_ctx ctx);

/* Text describing the last failure on the calling thread ("" if none). */
const char *pmlg_last_error(void);

/* C = A * B.  A is m x k (padded length lenA), B is k x n (lenB), C is m x n
 * with padded length lenA+lenB-1 (nothing is written when lenA or lenB is 0). */
pmlg_status pmlg_polmat_mul(pmlg_ctx ctx, uint64_t *C, const uint64_t *A, const uint64_t *B,
                            int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                            int method);
/* Same, on device pointers, enqueued on `stream` (a hipStream_t, or NULL for
 * the context's stream); returns without synchronising. */
pmlg_status pmlg_polmat_mul_dev(pmlg_ctx ctx, uint64_t *dC, const uint64_t *dA,
                                const uint64_t *dB, int64_t m, int64_t k, int64_t n,
                                int64_t lenA, int64_t lenB, int method, void *stream);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same product as pmlg_polmat_mul_dev, with identical
 * output, but every buffer beyond A, B and C (the evaluation arrays, and the
 * merge buffer of transforms above the LDS limit) is carved from a fixed
 * device array of PMLG_NOALLOC_SCRATCH_BYTES per GPU, held by the library
 * until the process ends; the call makes no hipMalloc, hipFree or other
 * allocating call.  Limits, each reported as an error before any work is
 * enqueued: the evaluations must fit the array ((m*k + k*n + m*n) * N words,
 * plus m*n*N when N exceeds 2^13 (64-bit words) / 2^14 (p < 2^30), N the
 * transform length 2^ceil(log2(lenA + lenB - 1)): PMLG_ERR_ARG); the product
 * must be evaluated modulo p itself, not by the multi-prime path
 * (PMLG_ERR_UNSUPPORTED); and the context's root tables for that transform
 * length must exist: any earlier product of that length on the context builds
 * them (PMLG_ERR_UNSUPPORTED otherwise).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (several streams, threads or
 * contexts) race on it and must be serialised by the caller.  No stage
 * timing (pmlg_profile_enable) is recorded for these calls. */
#define PMLG_NOALLOC_SCRATCH_BYTES (256ull << 20)
pmlg_status pmlg_polmat_mul_dev_noalloc(pmlg_ctx ctx, uint64_t *dC, const uint64_t *dA,
                                        const uint64_t *dB, int64_t m, int64_t k, int64_t n,
                                        int64_t lenA, int64_t lenB, int method, void *stream);

/* out = (A*B mod x^nhi) div x^nlo, padded length max(nhi-nlo, 0). */
pmlg_status pmlg_polmat_mulmid(pmlg_ctx ctx, uint64_t *out, const uint64_t *A,
                               const uint64_t *B, int64_t m, int64_t k, int64_t n,
                               int64_t lenA, int64_t lenB, int64_t nlo, int64_t nhi,
                               int method);
pmlg_status pmlg_polmat_mulmid_dev(pmlg_ctx ctx, uint64_t *dout, const uint64_t *dA,
                                   const uint64_t *dB, int64_t m, int64_t k, int64_t n,
                                   int64_t lenA, int64_t lenB, int64_t nlo, int64_t nhi,
                                   int method, void *stream);

/* Products of polynomial matrices held entry by entry in host memory, as
 * FLINT holds them (nmod_poly_mat_multiply, nmod_poly_mat_mulmid:
 * nmod_poly_mat_multiply.h:92,122-124).  Entry (i, j) of the r x c matrix X
 * has lenX[i*c + j] coefficients at X[i*c + j].  The coefficients are packed
 * straight into pinned staging chunks by several host threads and copied by
 * DMA while the next chunk is packed; results come back the same way.
 * out[i*n + j] must have room for nhi - nlo words (lenA + lenB - 1 for the
 * product, with lenA, lenB the largest entry lengths); lenOut receives each
 * output entry's length without trailing zeros, and only that many words are
 * written.  out must not overlap A or B. */
pmlg_status pmlg_polmat_mul_entries(pmlg_ctx ctx, uint64_t *const *C, int64_t *lenC,
                                    const uint64_t *const *A, const int64_t *lenA,
                                    const uint64_t *const *B, const int64_t *lenB, int64_t m,
                                    int64_t k, int64_t n, int method);
pmlg_status pmlg_polmat_mulmid_entries(pmlg_ctx ctx, uint64_t *const *out, int64_t *lenOut,
                                       const uint64_t *const *A, const int64_t *lenA,
                                       const uint64_t *const *B, const int64_t *lenB, int64_t m,
                                       int64_t k, int64_t n, int64_t nlo, int64_t nhi, int method);

/* Shift-ordered weak Popov approximant basis of F (m x n, padded length lenF)
 * at order `order` (same order for every column), by divide and conquer on the
 * order (nmod_poly_mat_pmbasis) or iteratively (nmod_poly_mat_mbasis).
 * P receives m x m entries of capacity capP >= order + 1 (entry (i,j)
 * coefficient d at P[(i*m + j)*capP + d]; every coefficient up to capP is
 * written, zeros past the basis); *lenP its normalised length;
 * shift (length m) is the input shift on entry and the shifted row degree of
 * P on return.  1 <= m <= 256, 1 <= n <= 128, with n (m | 1) words of the
 * transposed residual and the step's buffers within one workgroup's LDS
 * (PMLG_ERR_UNSUPPORTED otherwise; 160 x 80 fits, 256 x 62 with 64-bit words,
 * 256 x 128 for p < 2^30).  The products inside PM-Basis run on the device
 * with the given pointwise method.  pmlg_mbasis returns
 * PMLG_ERR_UNSUPPORTED when m^2 (order + 1) + m n order >= 2^31 (PM-Basis
 * has no such limit: its M-Basis leaves have order at most 256). */
pmlg_status pmlg_pmbasis(pmlg_ctx ctx, uint64_t *P, int64_t *lenP, int64_t *shift,
                         const uint64_t *F, int64_t m, int64_t n, int64_t lenF, int64_t order,
                         int64_t capP, int method);
pmlg_status pmlg_mbasis(pmlg_ctx ctx, uint64_t *P, int64_t *lenP, int64_t *shift,
                        const uint64_t *F, int64_t m, int64_t n, int64_t lenF, int64_t order,
                        int64_t capP);
/* PM-Basis with F and P in device memory (P capacity capP >= order + 1, its
 * length in *lenP is the computed one, not normalised: pmlg_pmbasis's basis
 * followed by zero coefficients); shift stays on the host.  Work is enqueued on
 * `stream` (NULL: the context's); the call reads one control word back per
 * M-Basis leaf and returns with the stream idle. */
pmlg_status pmlg_pmbasis_dev(pmlg_ctx ctx, uint64_t *dP, int64_t capP, int64_t *lenP,
                             int64_t *shift, const uint64_t *dF, int64_t m, int64_t n,
                             int64_t lenF, int64_t order, int method, void *stream);
/* M-Basis / PM-Basis crossover (PMBASIS_THRES, nmod_poly_mat_approximant.h:101,
 * a compile-time 32 in the reference).  The basis is the same for every leaf
 * order; only the time changes.  order 0 (the default): calls of order >= 512
 * use a leaf order measured once per context and (m, n, method) among 16, 32,
 * 64, 128 and 256 (two leaves of order t plus their products against one leaf
 * of order 2t, timed on seeded synthetic input), smaller calls use 32;
 * order in [1, 256]: that leaf order for every call. */
pmlg_status pmlg_pmbasis_set_leaf_order(pmlg_ctx ctx, int64_t order);
/* The leaf order a PM-Basis call of order >= 512 on an m x n input would use
 * (measured now if it is not yet known). */
pmlg_status pmlg_pmbasis_leaf_order(pmlg_ctx ctx, int64_t m, int64_t n, int method, int64_t *order);

/* Minimal left kernel bases (nmod_poly_mat_kernel.h:91-105): a basis of
 * {u : u F = 0} in shift-ordered weak Popov form (pivots increasing), for F
 * m x n with padded length lenF.
 *   algo PMLG_KERNEL_ZLS          replaces nmod_poly_mat_kernel_zls_approx
 *                                 (kernel.c:232-478, Zhou-Labahn-Storjohann);
 *   algo PMLG_KERNEL_VIA_APPROX   replaces nmod_poly_mat_kernel_via_approx
 *                                 (kernel.c:148-228, one approximant basis).
 * K receives m x m entries of capacity capK (entry (i,j) coefficient d at
 * K[(i*m + j)*capK + d]); its first *nullity rows form the basis, *lenK is
 * their normalised length.  pivind and shift have m entries: on return the
 * first *nullity hold the pivot indices and the shifted row degrees of the
 * basis rows (shift holds the input shift on entry).  Every approximant
 * basis, product and middle product runs on the device.  When capK < *lenK,
 * K is not written and PMLG_ERR_ARG is returned with *lenK and *nullity set:
 * call again with capK >= *lenK.  1 <= m <= 256; n <= 128 for
 * PMLG_KERNEL_VIA_APPROX; the modulus must be prime (p = 2 included). */
#define PMLG_KERNEL_ZLS 0
#define PMLG_KERNEL_VIA_APPROX 1
pmlg_status pmlg_kernel_basis(pmlg_ctx ctx, uint64_t *K, int64_t capK, int64_t *lenK,
                              int64_t *nullity, int64_t *pivind, int64_t *shift, const uint64_t *F,
                              int64_t m, int64_t n, int64_t lenF, int algo, int method);
/* Same with F and K in device memory, enqueued on `stream` (NULL: the
 * context's); synchronises it before returning (the recursion reads degrees
 * back at every level). */
pmlg_status pmlg_kernel_basis_dev(pmlg_ctx ctx, uint64_t *dK, int64_t capK, int64_t *lenK,
                                  int64_t *nullity, int64_t *pivind, int64_t *shift,
                                  const uint64_t *dF, int64_t m, int64_t n, int64_t lenF, int algo,
                                  int method, void *stream);

/* Rank (and determinant when r == c, else 0) of an r x c constant matrix
 * (row-major, host pointer, r*c <= 2^22), by Gaussian elimination on the
 * device.  Used by the approximant-basis certificate; no reference counterpart
 * (FLINT's nmod_mat_rank / nmod_mat_det there). */
pmlg_status pmlg_const_rank_det(pmlg_ctx ctx, const uint64_t *M, int64_t r, int64_t c,
                                int64_t *rank, uint64_t *det);
/* out[i][j] = A[i][j](x) for an r x c polynomial matrix (host pointers). */
pmlg_status pmlg_polmat_eval(pmlg_ctx ctx, uint64_t *out, const uint64_t *A, int64_t r, int64_t c,
                             int64_t len, uint64_t x);

/* ------------------------------------------------------------------------
 * Batched products (no reference counterpart: the reference multiplies one
 * pair per nmod_poly_mat_multiply call, mul.c:19-97).
 * ------------------------------------------------------------------------ */
typedef struct {
    pmlg_ctx ctx;       /* modulus of this product; every context of a batch on one device */
    const uint64_t *A;  /* device, m x k entries of padded length lenA */
    const uint64_t *B;  /* device, k x n entries of padded length lenB */
    uint64_t *C;        /* device, m x n entries of padded length lenA + lenB - 1 */
    int64_t m, k, n, lenA, lenB;
    int method;         /* PMLG_METHOD_*: AUTO = learned model, autotuned cache, then threshold */
    int config;         /* >= 0: force configuration `config` (0 .. pmlg_num_configs()-1) */
} pmlg_mul_desc;

typedef struct pmlg_batch_plan_s *pmlg_batch_plan;

#define PMLG_BATCH_GRAPH 1 /* capture the launches into a hipGraph on the first run */

/* Plans `count` independent products C_i = A_i * B_i (nothing is written for
 * a product with lenA or lenB 0).  The plan allocates the evaluation arena
 * and keeps the descriptors' device pointers: they must stay valid while the
 * plan is run.  Products are grouped so that one launch per stage serves all
 * products sharing a modulus and transform length. */
pmlg_status pmlg_batch_plan_create(pmlg_batch_plan *plan, const pmlg_mul_desc *descs,
                                   int64_t count, int flags);
/* Enqueues the whole batch on `stream` (NULL: the plan's stream); with
 * PMLG_BATCH_GRAPH it replays one captured graph.  Does not synchronise. */
pmlg_status pmlg_batch_plan_run(pmlg_batch_plan plan, void *stream);
/* Waits for the plan's own stream. */
pmlg_status pmlg_batch_plan_sync(pmlg_batch_plan plan);
/* Eager runs with HIP events between the stages on the plan's stream:
 * ms[0] forward transforms, ms[1] direct convolutions and pointwise
 * products, ms[2] inverse transforms (averages over `reps` runs). */
pmlg_status pmlg_batch_plan_profile(pmlg_batch_plan plan, int reps, double *ms);
/* Stage timing of graph replays: the plan's next run captures one graph per
 * stage (forward transforms | direct convolutions and pointwise products |
 * inverse transforms); later runs launch the three graphs in turn on the
 * stream with HIP events recorded before, between and after them, from
 * `ring` event sets used in turn.  ring 0 turns it off.  The plan needs
 * PMLG_BATCH_GRAPH. */
pmlg_status pmlg_batch_plan_time_stages(pmlg_batch_plan plan, int ring);
/* Waits for the timed replays and returns the stage times (ms[0..2])
 * averaged over each event set's last replay since the previous call, and
 * the number of replays averaged; the counts restart at zero. */
pmlg_status pmlg_batch_plan_stage_times(pmlg_batch_plan plan, double *ms, int64_t *replays);
/* Configuration chosen for product i (-1 for an empty product). */
pmlg_status pmlg_batch_plan_config(pmlg_batch_plan plan, int64_t i, int *config);
/* Kernel launches per run and arena bytes. */
pmlg_status pmlg_batch_plan_stats(pmlg_batch_plan plan, int64_t *launches, uint64_t *arena_bytes);
void pmlg_batch_plan_destroy(pmlg_batch_plan plan);
/* One-shot: plan, run, synchronise, destroy. */
pmlg_status pmlg_polmat_mul_batch(const pmlg_mul_desc *descs, int64_t count, void *stream);

/* n residues in [0, p) from a SplitMix64 stream of `seed`, written on the
 * device (synthetic inputs of large workloads without a host round trip). */
pmlg_status pmlg_fill_random(pmlg_ctx ctx, uint64_t *dst, int64_t n, uint64_t seed, void *stream);

/* Configurations: 0 direct convolution; 1..7 transforms with the plain or
 * MFMA pointwise kernel and different transform radices / MFMA block sizes. */
int pmlg_num_configs(void);
int pmlg_config_valid(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                      int config);
/* Configuration the dispatcher picks for this shape under `method`. */
int pmlg_choose_config(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                       int method);
/* Autotuner: times every valid configuration on device-generated inputs of
 * this shape (one warm-up and `reps` timed runs each; ms_per_config[c] = -1
 * for a configuration not timed), records the fastest for the shape class in
 * the context, and returns it in *best_config.  Each configuration runs a
 * plan of up to 16 identical products (enough for about 2^27 words of
 * evaluations, as in the shared launches of a real plan); the times are per
 * product. */
pmlg_status pmlg_autotune(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA,
                          int64_t lenB, int reps, double *ms_per_config, int *best_config);

/* ------------------------------------------------------------------------
 * Learned dispatch (no reference counterpart: the reference dispatches by
 * hand-set thresholds, mul.c:52-96 and PMBASIS_THRES).  A small fp32 MLP
 * maps the features (m/64, k/64, n/64, log2(lenA + lenB - 1)/16, bits(p)/64)
 * through three hidden ReLU layers of 64 to one predicted log(ms) per
 * configuration; training, gradients and inference run on the device.
 * ------------------------------------------------------------------------ */
#define PMLG_MLP_IN 5
#define PMLG_MLP_HIDDEN 64
#define PMLG_MLP_OUT 8 /* = pmlg_num_configs() */
#define PMLG_MLP_PARAMS 9224
typedef struct pmlg_mlp_s *pmlg_mlp;

/* He-uniform initial weights from `seed`, zero biases. */
pmlg_status pmlg_mlp_create(pmlg_mlp *mlp, int device, uint64_t seed);
void pmlg_mlp_destroy(pmlg_mlp mlp);
/* Flat parameters: for each layer l, W_l (in x out, row-major) then b_l. */
pmlg_status pmlg_mlp_get_params(pmlg_mlp mlp, float *params);
pmlg_status pmlg_mlp_set_params(pmlg_mlp mlp, const float *params);
/* out[r][c] = predicted log(ms) of configuration c for feature row r (host arrays). */
pmlg_status pmlg_mlp_forward(pmlg_mlp mlp, const float *X, int64_t rows, float *out);
/* Masked mean-squared error over the rows and its gradient (PMLG_MLP_PARAMS floats). */
pmlg_status pmlg_mlp_loss_grad(pmlg_mlp mlp, const float *X, const float *Y, const float *mask,
                               int64_t rows, float *loss, float *grads);
/* Adam on shuffled mini-batches (rows resident on the device); *final_loss =
 * masked MSE of the final parameters over the whole set. */
pmlg_status pmlg_mlp_train(pmlg_mlp mlp, const float *X, const float *Y, const float *mask,
                           int64_t rows, int epochs, int batch, float lr, uint64_t seed,
                           float *final_loss);
/* Feature row of a shape. */
void pmlg_mlp_features(int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB, uint64_t p,
                       float *x);
/* Model consulted by PMLG_METHOD_AUTO (NULL: none; the caller keeps it alive). */
pmlg_status pmlg_dispatch_set_model(pmlg_mlp mlp);
/* Measured refinement of the model's choice (on by default; PMLG_REFINE=0 in
 * the environment or this call turns it off): when configurations other
 * than the model's best are predicted within 15% of it, those candidates are
 * timed once per shape class, on replicated products as in pmlg_autotune,
 * and the fastest is used for every later product of the class. */
void pmlg_dispatch_set_refine(int on);

/* Per-stage device timing (HIP events around each launch).  Stages:
 * 0 forward NTT, 1 pointwise product, 2 inverse NTT (+ merge). */
void pmlg_profile_enable(pmlg_ctx ctx, int on);
/* Synchronises, then returns the accumulated milliseconds and launch counts
 * per stage (arrays of 3) and resets the accumulators. */
pmlg_status pmlg_profile_read(pmlg_ctx ctx, double *ms, int64_t *launches);

/* Plain device-to-device copy kernel (16 bytes per lane; pointers and size
 * 16-byte aligned), enqueued on `stream` or the context's stream.  Used as the
 * reference kernel when hardware byte counters are calibrated. */
pmlg_status pmlg_dev_copy(pmlg_ctx ctx, void *dst, const void *src, uint64_t bytes, void *stream);

/* Device memory helpers (plain hipMalloc/hipMemcpy on the context's device). */
pmlg_status pmlg_dev_alloc(pmlg_ctx ctx, void **ptr, uint64_t bytes);
void pmlg_dev_free(pmlg_ctx ctx, void *ptr);
pmlg_status pmlg_memcpy_h2d(pmlg_ctx ctx, void *dst, const void *src, uint64_t bytes);
pmlg_status pmlg_memcpy_d2h(pmlg_ctx ctx, void *dst, const void *src, uint64_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* PMLGPU_H */
