// Context object shared by the host translation units of libpmlgpu.
#pragma once

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pmlgpu.h"
#include "internal.hpp"

namespace pmlg {

struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t b)
    {
        if (b <= bytes)
            return hipSuccess;
        if (ptr) {
            // work enqueued earlier (possibly on a caller's stream) may still read it
            (void)hipDeviceSynchronize();
            (void)hipFree(ptr);
        }
        ptr = nullptr;
        bytes = 0;
        size_t want = b + b / 4;  // some headroom against regrowth
        hipError_t e = hipMalloc(&ptr, want);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        bytes = want;
        return hipSuccess;
    }
    void release()
    {
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

struct Tables {
    void *mem = nullptr;
    NttTab<uint64_t> t64{};
    NttTab<uint32_t> t32{};
    // long transforms (N above 4 times the LDS limit): column-pass tables;
    // the full tables (tw, twq, itw, itwq) are not built for them
    ColTab<uint64_t> c64{};
    ColTab<uint32_t> c32{};
};

// Pinned double-buffered staging chunks with their copy stream, one per
// host thread of the entry-wise transfers (staging.hip).
constexpr int kMaxStageThreads = 16;
int stage_threads();     // PMLG_STAGE_THREADS, default 2
size_t stage_bytes();    // PMLG_STAGE_MIB MiB per chunk, default 2
struct Stager {
    void *pin[2] = {nullptr, nullptr};
    size_t bytes = 0;
    hipStream_t s = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool used[2] = {false, false};
};

struct Prof {
    int stage;
    hipEvent_t a, b;
};

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess)
            prev = -1;
        if (prev != dev)
            (void)hipSetDevice(dev);
    }
    ~DeviceGuard()
    {
        if (prev >= 0)
            (void)hipSetDevice(prev);
    }
};

pmlg_status fail(pmlg_status st, const std::string &msg);
pmlg_status hip_fail(hipError_t e, const char *where);

#define HIP_TRY(expr, where)                \
    do {                                    \
        hipError_t _e = (expr);             \
        if (_e != hipSuccess)               \
            return hip_fail(_e, where);     \
    } while (0)

}  // namespace pmlg

struct pmlg_ctx_s {
    uint64_t p = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    pmlg::ModConst mc{};
    bool small = false;  // p < 2^30: 32-bit words on the device
    int max_log = 0;     // log2 of the largest supported transform
    uint64_t root = 0;   // primitive 2^max_log-th root of unity
    std::vector<pmlg::Tables> tabs;
    pmlg::DevBuf in_a, in_b, out_c, ev_a, ev_b, ev_c, tmp, arena;
    std::mutex mu;
    pmlg::NttTune tune;
    bool profile = false;
    std::vector<pmlg::Prof> pending;
    double ms[3] = {0, 0, 0};
    int64_t launches[3] = {0, 0, 0};
    // autotuner: best configuration per shape class, and its timing buffers
    std::mutex tune_mu;
    std::map<uint64_t, int> tuned;
    // measured refinement of the learned choice, per shape class
    std::map<uint64_t, int> refined;
    // PM-Basis leaf order (M-Basis / PM-Basis crossover): fixed when > 0,
    // else measured once per (m, n, method) and kept here
    int leaf_order = 0;
    std::map<uint64_t, int> leaf_orders;
    pmlg::DevBuf tune_io, tune_ws;
    // moduli without the roots a product needs: products modulo three FFT
    // primes (contexts created on first use) recombined by CRT
    bool is_prime = true;
    pmlg_ctx crt[3] = {nullptr, nullptr, nullptr};
    pmlg::DevBuf crt_in, crt_out;
    // the three products run on the sub-contexts' streams, forked from and
    // joined back to the caller's stream (events created on first use)
    hipEvent_t crt_fork = nullptr, crt_join[3] = {nullptr, nullptr, nullptr};
    // pinned staging for entry-wise host transfers (created on first use)
    std::vector<pmlg::Stager> stagers;
};

namespace pmlg {

int ceil_log2(int64_t x);
pmlg_status check_ctx(pmlg_ctx ctx);
hipError_t fill_zero(uint64_t *p, int64_t n, hipStream_t s);
// Coefficients [nlo, nhi) of A*B into dOut (padded length nhi - nlo), device
// pointers, operand strides = their padded lengths; ctx->mu must be held.
pmlg_status mulmid_dev_locked(pmlg_ctx ctx, uint64_t *dOut, const uint64_t *dA,
                              const uint64_t *dB, int64_t m, int64_t k, int64_t n, int64_t lenA,
                              int64_t lenB, int64_t nlo, int64_t nhi, int method, hipStream_t s);
// Coefficients [nlo, nhi) of A*B for a modulus whose roots of unity do not
// reach the transform length (or a composite modulus): the exact integer
// product modulo three FFT primes, recombined (Garner) and reduced mod p.
// Lengths already clipped as in mulmid_dev_strided; ctx->mu held.
pmlg_status crt_product(pmlg_ctx ctx, uint64_t *dOut, int64_t out_stride, const uint64_t *dA,
                        int64_t strideA, const uint64_t *dB, int64_t strideB, int64_t m, int64_t k,
                        int64_t n, int64_t la, int64_t lb, int64_t nlo, int64_t hi, int method,
                        hipStream_t s);
// Same with explicit operand strides (entry e of A at dA + e*strideA, length lenA <= strideA).
pmlg_status mulmid_dev_strided(pmlg_ctx ctx, uint64_t *dOut, int64_t out_stride,
                               const uint64_t *dA, int64_t strideA, const uint64_t *dB,
                               int64_t strideB, int64_t m, int64_t k, int64_t n, int64_t lenA,
                               int64_t lenB, int64_t nlo, int64_t nhi, int method, hipStream_t s);

// Approximant basis (PM-Basis, or M-Basis when iterative) of F (entry e at
// dF + e*strideF, length lenF) into dP (capacity capP >= order + 1) or, when hP
// is set, into host memory with a normalised length; shift in/out on the host.
// ctx->mu held.
pmlg_status basis_common(pmlg_ctx ctx, uint64_t *dP, int64_t capP, int64_t *lenP,
                         int64_t *shift, const uint64_t *dF, int64_t strideF, int64_t m, int64_t n,
                         int64_t lenF, int64_t order, int method, bool iterative, hipStream_t s,
                         uint64_t *hP);
// Rank of the r x c matrix of words at dM (row-major, overwritten), with
// dScratch holding r + 2 words; synchronises s.  ctx->mu held.
pmlg_status const_rank_dev(pmlg_ctx ctx, uint64_t *dM, uint64_t *dScratch, int64_t r, int64_t c,
                           int64_t *rank, hipStream_t s);

// Entry-wise host transfers through the pinned stagers (ctx->mu held).
// Upload: device entry e at dst + e*width <- ptr[e] (len[e] words, clipped and
// zero padded); returns when complete.  Download: ptr[e] <- device entry e
// of width words, normalised (lenOut[e] = length without trailing zeros),
// after the work enqueued on `after`; with zero_to > 0 each ptr[e] is also
// zero-filled from its length up to zero_to words.
pmlg_status stage_upload(pmlg_ctx ctx, uint64_t *dst, const uint64_t *const *ptr,
                         const int64_t *len, int64_t count, int64_t width);
pmlg_status stage_download(pmlg_ctx ctx, uint64_t *const *ptr, int64_t *lenOut,
                           const uint64_t *src, int64_t count, int64_t width, hipStream_t after,
                           int64_t zero_to = 0);
void stagers_release(pmlg_ctx ctx);

}  // namespace pmlg
