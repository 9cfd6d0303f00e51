// Host side of the device library: per-prime contexts, twiddle tables,
// scratch buffers, stage timing and the extern "C" status API of pmlgpu.h.
#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "ctx.hpp"
#include "host_arith.hpp"

using namespace pmlg;

namespace pmlg {

thread_local std::string g_last_error;

pmlg_status fail(pmlg_status st, const std::string &msg)
{
    g_last_error = msg;
    return st;
}

pmlg_status hip_fail(hipError_t e, const char *where)
{
    return fail(PMLG_ERR_DEVICE, std::string(where) + ": " + hipGetErrorString(e));
}





}  // namespace pmlg

namespace pmlg {


int ceil_log2(int64_t x)
{
    int l = 0;
    while (((int64_t)1 << l) < x)
        ++l;
    return l;
}

// Builds (once) the root tables for transform length 2^logN.  Up to 4 times
// the LDS limit: the full tables w^t, w^-t (t < N, with Shoup companions)
// then the per-stage tables of the LDS sub-transform.  Longer: the
// sub-transform's per-stage tables, then the column pass's tables (w_R^k and
// w_R^-k for k < R/2, and the two-level twist tables w^a, w^-a = lo[a mod S]
// * hi[a / S], S = 2^ceil(logN / 2)), so tables stay O(sqrt(N) + R + M).
pmlg_status get_tables(pmlg_ctx ctx, int logN, const Tables **out)
{
    if ((int)ctx->tabs.size() <= logN)
        ctx->tabs.resize(logN + 1);
    Tables &T = ctx->tabs[logN];
    if (!T.mem) {
        const int64_t N = (int64_t)1 << logN;
        const uint64_t p = ctx->p;
        const uint64_t w = host_powmod(ctx->root, (uint64_t)1 << (ctx->max_log - logN), p);
        const uint64_t wi = host_invmod(w, p);
        const size_t wsz = ctx->small ? 4 : 8;
        const int lim = ctx->small ? NttLimits<uint32_t>::max_log_m : NttLimits<uint64_t>::max_log_m;
        const int logM = logN < lim ? logN : lim;
        const int64_t M = (int64_t)1 << logM;
        const bool longN = logN > lim + 2;
        const int logR = logN - logM;
        const int64_t R = (int64_t)1 << logR;
        const int slog = (logN + 1) / 2;
        const int64_t S = (int64_t)1 << slog, H = N >> slog;
        // word offsets: [full 4N | stage 4M] or [stage 4M | rt 4 (R/2) | lo, hi (4 (S + H))]
        const int64_t full = longN ? 0 : 4 * N;
        const int64_t stage0 = full;
        const int64_t rt0 = stage0 + 4 * M;
        const int64_t tw0 = rt0 + (longN ? 2 * R : 0);
        const int64_t words = tw0 + (longN ? 4 * (S + H) : 0);
        std::vector<unsigned char> h((size_t)words * wsz);
        auto put = [&](int64_t at, uint64_t v) {
            if (ctx->small)
                reinterpret_cast<uint32_t *>(h.data())[at] = (uint32_t)v;
            else
                reinterpret_cast<uint64_t *>(h.data())[at] = v;
        };
        auto pre = [&](uint64_t v) {
            return ctx->small ? host_shoup_pre32(v, p) : host_shoup_pre64(v, p);
        };
        // value and companion of x^j, j < cnt, at [at, at + cnt) and [at + cnt, at + 2 cnt)
        auto powers = [&](int64_t at, int64_t cnt, uint64_t x) {
            uint64_t v = 1 % p;
            for (int64_t j = 0; j < cnt; ++j) {
                put(at + j, v);
                put(at + cnt + j, pre(v));
                v = host_mulmod(v, x, p);
            }
        };
        if (!longN) {
            powers(0, N, w);       // tw, twq
            powers(2 * N, N, wi);  // itw, itwq
        }
        {
            // sub-transform root w_M = w^(N/M); stage s entry x = w_M^(x 2^s)
            const uint64_t wm = host_powmod(w, (uint64_t)(N / M), p);
            const uint64_t wmi = host_invmod(wm, p);
            for (int s = 0; s < logM; ++s) {
                const int64_t off = M - ((int64_t)1 << (logM - s));
                const int64_t cnt = (int64_t)1 << (logM - 1 - s);
                const uint64_t step = host_powmod(wm, (uint64_t)1 << s, p);
                const uint64_t stepi = host_powmod(wmi, (uint64_t)1 << s, p);
                uint64_t a = 1 % p, b = 1 % p;
                for (int64_t x = 0; x < cnt; ++x) {
                    const int64_t f = stage0 + 2 * (off + x);  // forward pair index (words)
                    const int64_t g = stage0 + 2 * M + 2 * (off + x);
                    put(f, a);
                    put(f + 1, pre(a));
                    put(g, b);
                    put(g + 1, pre(b));
                    a = host_mulmod(a, step, p);
                    b = host_mulmod(b, stepi, p);
                }
            }
        }
        if (longN) {
            const uint64_t wr = host_powmod(w, (uint64_t)M, p), wri = host_invmod(wr, p);
            powers(rt0, R / 2, wr);          // rt, rtq
            powers(rt0 + R, R / 2, wri);     // rti, rtiq
            powers(tw0, S, w);               // lo, loq
            powers(tw0 + 2 * S, H, host_powmod(w, (uint64_t)S, p));     // hi, hiq
            powers(tw0 + 2 * S + 2 * H, S, wi);                         // loi, loiq
            powers(tw0 + 4 * S + 2 * H, H, host_powmod(wi, (uint64_t)S, p));  // hii, hiiq
        }
        void *mem = nullptr;
        HIP_TRY(hipMalloc(&mem, h.size()), "hipMalloc(twiddles)");
        hipError_t e = hipMemcpy(mem, h.data(), h.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(mem);
            return hip_fail(e, "hipMemcpy(twiddles)");
        }
        T.mem = mem;
        const uint64_t ninv = host_invmod((uint64_t)(N % p), p);
        const uint64_t ninvr = host_mulmod(ninv, ctx->mc.r64, p);
        const uint64_t ninv2 = host_mulmod(ninv, 2 % p, p), ninvr2 = host_mulmod(ninvr, 2 % p, p);
        // w^(N/4) and w^(N/4) / 2 (p odd: transforms need a prime p >= 3)
        const uint64_t tquart = N >= 4 ? host_powmod(w, (uint64_t)(N / 4), p) : 0;
        const uint64_t thalf = host_mulmod(tquart, (p + 1) / 2, p);
        auto fill = [&](auto *a, auto &tab, auto &col) {
            using Wt = std::remove_const_t<std::remove_pointer_t<decltype(a)>>;
            const Wt *full_tw = longN ? nullptr : a;
            tab.tw = full_tw;
            tab.twq = longN ? nullptr : a + N;
            tab.itw = longN ? nullptr : a + 2 * N;
            tab.itwq = longN ? nullptr : a + 3 * N;
            tab.ninv = (Wt)ninv;
            tab.ninvq = (Wt)pre(ninv);
            tab.stw = a + stage0;
            tab.sitw = a + stage0 + 2 * M;
            tab.ninvr = (Wt)ninvr;
            tab.ninvrq = (Wt)pre(ninvr);
            tab.ninv2 = (Wt)ninv2;
            tab.ninv2q = (Wt)pre(ninv2);
            tab.ninvr2 = (Wt)ninvr2;
            tab.ninvr2q = (Wt)pre(ninvr2);
            tab.thalf = (Wt)thalf;
            tab.thalfq = (Wt)pre(thalf);
            tab.tquart = (Wt)tquart;
            tab.tquartq = (Wt)pre(tquart);
            if (longN) {
                col.rt = a + rt0;
                col.rtq = a + rt0 + R / 2;
                col.rti = a + rt0 + R;
                col.rtiq = a + rt0 + R + R / 2;
                col.lo = a + tw0;
                col.loq = a + tw0 + S;
                col.hi = a + tw0 + 2 * S;
                col.hiq = a + tw0 + 2 * S + H;
                col.loi = a + tw0 + 2 * S + 2 * H;
                col.loiq = a + tw0 + 3 * S + 2 * H;
                col.hii = a + tw0 + 4 * S + 2 * H;
                col.hiiq = a + tw0 + 4 * S + 3 * H;
                col.slog = slog;
            }
        };
        if (ctx->small)
            fill(reinterpret_cast<const uint32_t *>(mem), T.t32, T.c32);
        else
            fill(reinterpret_cast<const uint64_t *>(mem), T.t64, T.c64);
    }
    *out = &T;
    return PMLG_OK;
}

void prof_begin(pmlg_ctx ctx, int stage, hipStream_t s, Prof *pr)
{
    pr->stage = -1;
    if (!ctx->profile)
        return;
    if (hipEventCreate(&pr->a) != hipSuccess)
        return;
    if (hipEventCreate(&pr->b) != hipSuccess) {
        (void)hipEventDestroy(pr->a);
        return;
    }
    pr->stage = stage;
    (void)hipEventRecord(pr->a, s);
}

void prof_end(pmlg_ctx ctx, hipStream_t s, Prof *pr)
{
    if (pr->stage < 0)
        return;
    (void)hipEventRecord(pr->b, s);
    ctx->pending.push_back(*pr);
}

// Fill kernel (used instead of hipMemsetAsync so that the sequence stays
// capturable in a HIP graph).
__global__ void fill_zero_kernel(uint64_t *__restrict__ p, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// Plain 16-byte-per-lane device copy: the reference point for byte counters.
__global__ void copy16_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, int64_t n16)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

hipError_t fill_zero(uint64_t *p, int64_t n, hipStream_t s)
{
    if (n <= 0)
        return hipSuccess;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096)
        blocks = 4096;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

template <typename W>
const NttTab<W> &tab_of(const Tables &T);
template <>
const NttTab<uint64_t> &tab_of<uint64_t>(const Tables &T)
{
    return T.t64;
}
template <>
const NttTab<uint32_t> &tab_of<uint32_t>(const Tables &T)
{
    return T.t32;
}

template <typename W>
const ColTab<W> &col_of(const Tables &T);
template <>
const ColTab<uint64_t> &col_of<uint64_t>(const Tables &T)
{
    return T.c64;
}
template <>
const ColTab<uint32_t> &col_of<uint32_t>(const Tables &T)
{
    return T.c32;
}

// Bytes of the transforms' work buffer for an m x k by k x n product at
// length 2^logN (0 when N fits 4 times the LDS limit... below it): the
// inverse's m*n*N words of W above the LDS limit, and the forward column
// pass's max(m*k, k*n)*N 64-bit words for long transforms.
template <typename W>
int64_t product_tmp_bytes(int64_t m, int64_t k, int64_t n, int logN)
{
    const int lim = NttLimits<W>::max_log_m;
    if (logN <= lim)
        return 0;
    const int64_t N = (int64_t)1 << logN;
    int64_t b = m * n * N * (int64_t)sizeof(W);
    if (logN > lim + 2)
        b = std::max<int64_t>(b, std::max(m * k, k * n) * N * 8);
    return b;
}

// Evaluation/interpolation pipeline for coefficients [lo, hi) of A*B, with a
// transform length 2^logN that is at least hi and at least lenC - lo, on the
// given root tables and evaluation buffers (Ah: m*k*N words of W, Bh: k*n*N,
// Ch: m*n*N, tmp: m*n*N when N exceeds the LDS limit, else unused).  Stage
// timing (HIP events) only when `prof` is set.
template <typename W>
pmlg_status run_product_on(pmlg_ctx ctx, const Tables *T, W *Ah, W *Bh, W *Ch, void *tmpv, bool prof,
                           uint64_t *dOut, int64_t out_stride, const uint64_t *dA,
                           const uint64_t *dB, int64_t m, int64_t k, int64_t n, int64_t strideA,
                           int64_t strideB, int64_t la, int64_t lb, int64_t lo, int64_t hi,
                           int logN, PointwiseKind kind, hipStream_t s)
{
    const NttTab<W> &tab = tab_of<W>(*T);
    const ColTab<W> *ct = &col_of<W>(*T);
    W *tmp = static_cast<W *>(tmpv);                   // inverse work buffer
    uint64_t *tmpf = static_cast<uint64_t *>(tmpv);    // forward column pass
    const int64_t N = (int64_t)1 << logN;
    const W p = (W)ctx->p;
    const int logP = eval_log_run_in<W>(logN), logPc = eval_log_run_out<W>(logN);
    // a product of length <= 3N/4 (no wrap-around) needs 3N/4 evaluations only
    const bool trunc = trunc_ok<W>(logN, la + lb - 1) && hi <= 3 * (N >> 2);
    const int64_t npts = trunc ? 3 * (N >> 2) : N;
    const NttTune &tune = ctx->tune;
    Prof pr;
    pr.stage = -1;
    if (prof)
        prof_begin(ctx, 0, s, &pr);
    HIP_TRY((launch_ntt_fwd<uint64_t, W>(dA, strideA, la, Ah, m * k, logN, logP, tune, tab, p, s,
                                         tmpf, ct, trunc)),
            "ntt_fwd(A)");
    if (prof) {
        prof_end(ctx, s, &pr);
        prof_begin(ctx, 0, s, &pr);
    }
    HIP_TRY((launch_ntt_fwd<uint64_t, W>(dB, strideB, lb, Bh, k * n, logN, logP, tune, tab, p, s,
                                         tmpf, ct, trunc)),
            "ntt_fwd(B)");
    if (prof)
        prof_end(ctx, s, &pr);
    // The 8-limb MFMA product leaves V 2^-64 mod p (Montgomery REDC) in [0, 4p);
    // the inverse transform then scales by N^-1 2^64 instead of N^-1.
    ModConst mc = ctx->mc;
    NttTab<W> itab = tab;
    if (kind == PW_MFMA && sizeof(W) == 8) {
        const unsigned __int128 vmax = (unsigned __int128)k * (ctx->p - 1) * (ctx->p - 1);
        mc.redc_small = vmax < (((unsigned __int128)(3 * ctx->p - 1)) << 64) ? 1 : 0;
        itab.ninv = tab.ninvr;
        itab.ninvq = tab.ninvrq;
        itab.ninv2 = tab.ninvr2;
        itab.ninv2q = tab.ninvr2q;
    }
    if (prof)
        prof_begin(ctx, 1, s, &pr);
    HIP_TRY(launch_pointwise<W>(kind, Ah, Bh, Ch, (int)m, (int)k, (int)n, npts, logP, logPc, mc, s),
            "pointwise");
    if (prof) {
        prof_end(ctx, s, &pr);
        prof_begin(ctx, 2, s, &pr);
    }
    HIP_TRY((launch_ntt_inv<W, uint64_t>(Ch, dOut, out_stride, lo, hi, m * n, logN, logPc, tune,
                                         itab, p, tmp, s, ct, trunc)),
            "ntt_inv");
    if (prof)
        prof_end(ctx, s, &pr);
    return PMLG_OK;
}

// The same on the context's tables and evaluation buffers (built or grown here).
template <typename W>
pmlg_status run_product(pmlg_ctx ctx, uint64_t *dOut, int64_t out_stride, const uint64_t *dA,
                        const uint64_t *dB, int64_t m, int64_t k, int64_t n, int64_t strideA,
                        int64_t strideB, int64_t la, int64_t lb, int64_t lo, int64_t hi, int logN,
                        PointwiseKind kind, hipStream_t s)
{
    const Tables *T = nullptr;
    pmlg_status st = get_tables(ctx, logN, &T);
    if (st != PMLG_OK)
        return st;
    const int64_t N = (int64_t)1 << logN;
    HIP_TRY(ctx->ev_a.ensure((size_t)(m * k * N) * sizeof(W)), "hipMalloc(eval A)");
    HIP_TRY(ctx->ev_b.ensure((size_t)(k * n * N) * sizeof(W)), "hipMalloc(eval B)");
    HIP_TRY(ctx->ev_c.ensure((size_t)(m * n * N) * sizeof(W)), "hipMalloc(eval C)");
    void *tmp = nullptr;
    if (const int64_t tb = product_tmp_bytes<W>(m, k, n, logN)) {
        HIP_TRY(ctx->tmp.ensure((size_t)tb), "hipMalloc(tmp)");
        tmp = ctx->tmp.ptr;
    }
    return run_product_on<W>(ctx, T, static_cast<W *>(ctx->ev_a.ptr), static_cast<W *>(ctx->ev_b.ptr),
                             static_cast<W *>(ctx->ev_c.ptr), tmp, ctx->profile, dOut, out_stride,
                             dA, dB, m, k, n, strideA, strideB, la, lb, lo, hi, logN, kind, s);
}

pmlg_status check_ctx(pmlg_ctx ctx)
{
    if (!ctx)
        return fail(PMLG_ERR_ARG, "null context");
    return PMLG_OK;
}

pmlg_status pick_kind(pmlg_ctx ctx, int method, int64_t m, int64_t k, int64_t n,
                      PointwiseKind *kind)
{
    const bool mfma_ok = ctx->mc.limbs > 0 && mfma_supported((int)m, (int)k, (int)n, ctx->mc.limbs, ctx->p);
    if (method == PMLG_METHOD_PLAIN) {
        *kind = PW_PLAIN;
    } else if (method == PMLG_METHOD_MFMA) {
        if (!mfma_ok)
            return fail(PMLG_ERR_UNSUPPORTED,
                        "MFMA pointwise kernel: shape or modulus outside its limits");
        *kind = PW_MFMA;
    } else if (method == PMLG_METHOD_AUTO) {
        *kind = mfma_ok ? PW_MFMA : PW_PLAIN;
    } else {
        return fail(PMLG_ERR_ARG, "unknown method");
    }
    return PMLG_OK;
}

// Core of every product call: coefficients [nlo, nhi) of A*B into dOut
// (entry e at dOut + e*out_stride, out_stride >= nhi - nlo), all pointers on
// the device; entry e of A at dA + e*strideA with lenA <= strideA (same for B).
pmlg_status mulmid_dev_strided(pmlg_ctx ctx, uint64_t *dOut, int64_t out_stride,
                               const uint64_t *dA, int64_t strideA, const uint64_t *dB,
                               int64_t strideB, int64_t m, int64_t k, int64_t n, int64_t lenA,
                               int64_t lenB, int64_t nlo, int64_t nhi, int method, hipStream_t s)
{
    if (m < 0 || k < 0 || n < 0 || lenA < 0 || lenB < 0 || nlo < 0)
        return fail(PMLG_ERR_ARG, "negative dimension, length or nlo");
    if (m > INT32_MAX / 64 || k > INT32_MAX / 64 || n > INT32_MAX / 64)
        return fail(PMLG_ERR_ARG, "dimension too large");
    const int64_t width = nhi - nlo;
    if (width <= 0 || m == 0 || n == 0)
        return PMLG_OK;
    if (out_stride < width || strideA < lenA || strideB < lenB)
        return fail(PMLG_ERR_ARG, "stride shorter than the length it holds");
    if ((!dOut) || ((k > 0 && lenA > 0 && lenB > 0) && (!dA || !dB)))
        return fail(PMLG_ERR_ARG, "null buffer");
    const int64_t lenC = (lenA > 0 && lenB > 0) ? lenA + lenB - 1 : 0;
    const int64_t hi_eff = nhi < lenC ? nhi : lenC;
    if (k == 0 || hi_eff <= nlo) {
        HIP_TRY(fill_zero(dOut, m * n * out_stride, s), "fill_zero");
        return PMLG_OK;
    }
    // Only coefficients below hi_eff of each operand matter; they are read in
    // place with their original strides, so clip the lengths.
    const int64_t la = lenA < hi_eff ? lenA : hi_eff;
    const int64_t lb = lenB < hi_eff ? lenB : hi_eff;
    const int64_t lc = la + lb - 1;
    // cyclic length: no wrapped term may land in [nlo, hi_eff)
    int64_t need = hi_eff > lc - nlo ? hi_eff : lc - nlo;
    const int logN = ceil_log2(need);
    if (logN > ctx->max_log) {
        // no 2^logN-th root of unity mod p (or p composite): multi-prime path
        if (hi_eff < nhi)
            HIP_TRY(fill_zero(dOut, m * n * out_stride, s), "fill_zero");
        return crt_product(ctx, dOut, out_stride, dA, strideA, dB, strideB, m, k, n, la, lb, nlo,
                           hi_eff, method, s);
    }
    PointwiseKind kind;
    pmlg_status st = pick_kind(ctx, method, m, k, n, &kind);
    if (st != PMLG_OK)
        return st;
    if (hi_eff < nhi)  // coefficients beyond the product length are zero
        HIP_TRY(fill_zero(dOut, m * n * out_stride, s), "fill_zero");
    if (ctx->small)
        return run_product<uint32_t>(ctx, dOut, out_stride, dA, dB, m, k, n, strideA, strideB, la,
                                     lb, nlo, hi_eff, logN, kind, s);
    return run_product<uint64_t>(ctx, dOut, out_stride, dA, dB, m, k, n, strideA, strideB, la, lb,
                                 nlo, hi_eff, logN, kind, s);
}

pmlg_status mulmid_dev_locked(pmlg_ctx ctx, uint64_t *dOut, const uint64_t *dA,
                              const uint64_t *dB, int64_t m, int64_t k, int64_t n, int64_t lenA,
                              int64_t lenB, int64_t nlo, int64_t nhi, int method, hipStream_t s)
{
    const int64_t width = nhi - nlo;
    return mulmid_dev_strided(ctx, dOut, width > 0 ? width : 0, dA, lenA, dB, lenB, m, k, n, lenA,
                              lenB, nlo, nhi, method, s);
}

}  // namespace pmlg

// ---------------------------------------------------------------------------
// Fixed device scratch of pmlg_polmat_mul_dev_noalloc: the evaluation buffers
// of one product, in a file-scope device array instead of allocations.  Each
// device the library runs on holds its own copy for the life of the process.
// Host code never uses the array's name as an address: its device address is
// looked up once per device (hipGetSymbolAddress, with that device current)
// and kept in g_scratch_ptr.
// ---------------------------------------------------------------------------
static __device__ __align__(256) uint64_t scratch[PMLG_NOALLOC_SCRATCH_BYTES / 8];

namespace pmlg {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// device address of this device's scratch array (the device is current)
pmlg_status scratch_address(int device, void **out)
{
    if (device < 0 || device >= kMaxDevices)
        return fail(PMLG_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol,
                            "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return PMLG_OK;
}

constexpr int64_t align256(int64_t b)
{
    return (b + 255) & ~(int64_t)255;
}

// C = A * B with every buffer beyond A, B and C carved from the scratch
// array; ctx->mu held, ctx->device current.
template <typename W>
pmlg_status mul_noalloc_locked(pmlg_ctx ctx, uint64_t *dC, const uint64_t *dA, const uint64_t *dB,
                               int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                               int logN, PointwiseKind kind, hipStream_t s)
{
    const int64_t N = (int64_t)1 << logN;
    const int64_t bA = align256(m * k * N * (int64_t)sizeof(W));
    const int64_t bB = align256(k * n * N * (int64_t)sizeof(W));
    const int64_t bC = align256(m * n * N * (int64_t)sizeof(W));
    const int64_t bT = align256(product_tmp_bytes<W>(m, k, n, logN));
    const int64_t need = bA + bB + bC + bT;
    if (need > (int64_t)PMLG_NOALLOC_SCRATCH_BYTES)
        return fail(PMLG_ERR_ARG, "pmlg_polmat_mul_dev_noalloc: this product needs " +
                                      std::to_string(need) + " bytes of evaluation scratch, more "
                                      "than the fixed " +
                                      std::to_string((int64_t)PMLG_NOALLOC_SCRATCH_BYTES) +
                                      " bytes; use pmlg_polmat_mul_dev");
    void *base = nullptr;
    pmlg_status st = scratch_address(ctx->device, &base);
    if (st != PMLG_OK)
        return st;
    unsigned char *b = static_cast<unsigned char *>(base);
    W *Ah = reinterpret_cast<W *>(b);
    W *Bh = reinterpret_cast<W *>(b + bA);
    W *Ch = reinterpret_cast<W *>(b + bA + bB);
    void *tmp = bT ? static_cast<void *>(b + bA + bB + bC) : nullptr;
    const int64_t lenC = lenA + lenB - 1;
    return run_product_on<W>(ctx, &ctx->tabs[logN], Ah, Bh, Ch, tmp, false, dC, lenC, dA, dB, m,
                             k, n, lenA, lenB, lenA, lenB, 0, lenC, logN, kind, s);
}

}  // namespace
}  // namespace pmlg

// ---------------------------------------------------------------------------
// extern "C" API
// ---------------------------------------------------------------------------
extern "C" {

const char *pmlg_last_error(void)
{
    return g_last_error.c_str();
}

pmlg_status pmlg_ctx_create(pmlg_ctx *out, uint64_t p, int device)
{
    if (!out)
        return fail(PMLG_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (p < 2)
        return fail(PMLG_ERR_UNSUPPORTED, "modulus must be at least 2");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(PMLG_ERR_DEVICE, "no HIP device available");
    if (device < 0 || device >= ndev)
        return fail(PMLG_ERR_ARG, "device index out of range");
    pmlg_ctx ctx = new (std::nothrow) pmlg_ctx_s();
    if (!ctx)
        return fail(PMLG_ERR_ALLOC, "out of host memory");
    ctx->p = p;
    ctx->device = device;
    ctx->small = p < ((uint64_t)1 << 30);
    ctx->is_prime = p > 2 && host_is_prime(p);
    // the transforms keep values below 4p in 64-bit words: p >= 2^62 takes
    // the multi-prime path for every product
    const bool transforms = ctx->is_prime && p < ((uint64_t)1 << 62);
    if (transforms) {
        int v2 = __builtin_ctzll(p - 1);
        const int cap = ctx->small ? NttLimits<uint32_t>::max_log_n
                                   : NttLimits<uint64_t>::max_log_n;
        ctx->max_log = v2 < cap ? v2 : cap;
        ctx->root = host_root_of_unity_2k(p, ctx->max_log);
    } else {
        // no transform modulo p itself (composite, or p >= 2^62): every product
        // takes the multi-prime path
        ctx->max_log = -1;
        ctx->root = 0;
    }
    ctx->mc = host_modconst(p);
    // radix of the register transforms (autotuned values may replace these)
    if (const char *v = std::getenv("PMLG_NTT_FWD_LOGE"))
        ctx->tune.fwd_log_e = std::atoi(v) == 4 ? 4 : 5;
    if (const char *v = std::getenv("PMLG_NTT_INV_LOGE"))
        ctx->tune.inv_log_e = std::atoi(v) == 5 ? 5 : 4;
    DeviceGuard g(device);
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return hip_fail(e, "hipStreamCreate");
    }
    *out = ctx;
    return PMLG_OK;
}

void pmlg_ctx_destroy(pmlg_ctx ctx)
{
    if (!ctx)
        return;
    for (pmlg_ctx &c : ctx->crt)
        if (c) {
            pmlg_ctx_destroy(c);
            c = nullptr;
        }
    {
        DeviceGuard g(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (auto &pr : ctx->pending) {
            (void)hipEventDestroy(pr.a);
            (void)hipEventDestroy(pr.b);
        }
        for (auto &t : ctx->tabs)
            if (t.mem)
                (void)hipFree(t.mem);
        ctx->in_a.release();
        ctx->in_b.release();
        ctx->out_c.release();
        ctx->ev_a.release();
        ctx->ev_b.release();
        ctx->ev_c.release();
        ctx->tmp.release();
        ctx->arena.release();
        ctx->tune_io.release();
        ctx->tune_ws.release();
        ctx->crt_in.release();
        ctx->crt_out.release();
        if (ctx->crt_fork)
            (void)hipEventDestroy(ctx->crt_fork);
        for (hipEvent_t e : ctx->crt_join)
            if (e)
                (void)hipEventDestroy(e);
        stagers_release(ctx);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

int pmlg_ctx_max_log_len(pmlg_ctx ctx)
{
    return ctx ? ctx->max_log : -1;
}

void *pmlg_ctx_stream(pmlg_ctx ctx)
{
    return ctx ? (void *)ctx->stream : nullptr;
}

pmlg_status pmlg_ctx_sync(pmlg_ctx ctx)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    DeviceGuard g(ctx->device);
    HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    return PMLG_OK;
}

pmlg_status pmlg_polmat_mulmid_dev(pmlg_ctx ctx, uint64_t *dout, const uint64_t *dA,
                                   const uint64_t *dB, int64_t m, int64_t k, int64_t n,
                                   int64_t lenA, int64_t lenB, int64_t nlo, int64_t nhi,
                                   int method, void *stream)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    return mulmid_dev_locked(ctx, dout, dA, dB, m, k, n, lenA, lenB, nlo, nhi, method, s);
}

pmlg_status pmlg_polmat_mul_dev(pmlg_ctx ctx, uint64_t *dC, const uint64_t *dA,
                                const uint64_t *dB, int64_t m, int64_t k, int64_t n,
                                int64_t lenA, int64_t lenB, int method, void *stream)
{
    if (lenA <= 0 || lenB <= 0)
        return lenA < 0 || lenB < 0 ? fail(PMLG_ERR_ARG, "negative length") : PMLG_OK;
    return pmlg_polmat_mulmid_dev(ctx, dC, dA, dB, m, k, n, lenA, lenB, 0, lenA + lenB - 1,
                                  method, stream);
}

pmlg_status pmlg_polmat_mul_dev_noalloc(pmlg_ctx ctx, uint64_t *dC, const uint64_t *dA,
                                        const uint64_t *dB, int64_t m, int64_t k, int64_t n,
                                        int64_t lenA, int64_t lenB, int method, void *stream)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    if (m < 0 || k < 0 || n < 0 || lenA < 0 || lenB < 0)
        return fail(PMLG_ERR_ARG, "negative dimension or length");
    if (m > INT32_MAX / 64 || k > INT32_MAX / 64 || n > INT32_MAX / 64)
        return fail(PMLG_ERR_ARG, "dimension too large");
    if (lenA == 0 || lenB == 0 || m == 0 || n == 0)
        return PMLG_OK;
    if (!dC || (k > 0 && (!dA || !dB)))
        return fail(PMLG_ERR_ARG, "null buffer");
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    const int64_t lenC = lenA + lenB - 1;
    if (k == 0) {
        HIP_TRY(fill_zero(dC, m * n * lenC, s), "fill_zero");
        return PMLG_OK;
    }
    const int logN = ceil_log2(lenC);
    if (logN > ctx->max_log)
        return fail(PMLG_ERR_UNSUPPORTED,
                    "pmlg_polmat_mul_dev_noalloc: transform length 2^" + std::to_string(logN) +
                        " needs the multi-prime path, which allocates; use pmlg_polmat_mul_dev");
    if ((int)ctx->tabs.size() <= logN || !ctx->tabs[logN].mem)
        return fail(PMLG_ERR_UNSUPPORTED,
                    "pmlg_polmat_mul_dev_noalloc: the root tables of transform length 2^" +
                        std::to_string(logN) +
                        " are not built on this context; any earlier product of that transform "
                        "length on it (e.g. one pmlg_polmat_mul_dev call) builds them");
    PointwiseKind kind;
    st = pick_kind(ctx, method, m, k, n, &kind);
    if (st != PMLG_OK)
        return st;
    if (ctx->small)
        return mul_noalloc_locked<uint32_t>(ctx, dC, dA, dB, m, k, n, lenA, lenB, logN, kind, s);
    return mul_noalloc_locked<uint64_t>(ctx, dC, dA, dB, m, k, n, lenA, lenB, logN, kind, s);
}

pmlg_status pmlg_polmat_mulmid(pmlg_ctx ctx, uint64_t *out, const uint64_t *A, const uint64_t *B,
                               int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                               int64_t nlo, int64_t nhi, int method)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    if (m < 0 || k < 0 || n < 0 || lenA < 0 || lenB < 0 || nlo < 0)
        return fail(PMLG_ERR_ARG, "negative dimension, length or nlo");
    const int64_t width = nhi - nlo;
    if (width <= 0 || m == 0 || n == 0)
        return PMLG_OK;
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    hipStream_t s = ctx->stream;
    const size_t ba = (size_t)(m * k * lenA) * 8, bb = (size_t)(k * n * lenB) * 8;
    const size_t bc = (size_t)(m * n * width) * 8;
    HIP_TRY(ctx->in_a.ensure(ba ? ba : 8), "hipMalloc(A)");
    HIP_TRY(ctx->in_b.ensure(bb ? bb : 8), "hipMalloc(B)");
    HIP_TRY(ctx->out_c.ensure(bc), "hipMalloc(C)");
    // operands and result move through the pinned staging chunks (packed
    // host arrays seen as entries of one length each)
    auto upload = [&](void *dst, const uint64_t *src, int64_t count, int64_t len) -> pmlg_status {
        std::vector<const uint64_t *> ptr((size_t)count);
        std::vector<int64_t> lens((size_t)count, len);
        for (int64_t e = 0; e < count; ++e)
            ptr[(size_t)e] = src + e * len;
        return stage_upload(ctx, static_cast<uint64_t *>(dst), ptr.data(), lens.data(), count, len);
    };
    if (ba && (st = upload(ctx->in_a.ptr, A, m * k, lenA)) != PMLG_OK)
        return st;
    if (bb && (st = upload(ctx->in_b.ptr, B, k * n, lenB)) != PMLG_OK)
        return st;
    st = mulmid_dev_locked(ctx, static_cast<uint64_t *>(ctx->out_c.ptr),
                           static_cast<const uint64_t *>(ctx->in_a.ptr),
                           static_cast<const uint64_t *>(ctx->in_b.ptr), m, k, n, lenA, lenB, nlo,
                           nhi, method, s);
    if (st != PMLG_OK) {
        (void)hipStreamSynchronize(s);
        return st;
    }
    // every output word is written: each entry's zero top is filled
    std::vector<uint64_t *> po((size_t)(m * n));
    std::vector<int64_t> lo((size_t)(m * n));
    for (int64_t e = 0; e < m * n; ++e)
        po[(size_t)e] = out + e * width;
    st = stage_download(ctx, po.data(), lo.data(), static_cast<const uint64_t *>(ctx->out_c.ptr), m * n,
                        width, s, width);
    if (st != PMLG_OK) {
        (void)hipStreamSynchronize(s);
        return st;
    }
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    return PMLG_OK;
}

pmlg_status pmlg_polmat_mul(pmlg_ctx ctx, uint64_t *C, const uint64_t *A, const uint64_t *B,
                            int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                            int method)
{
    if (lenA <= 0 || lenB <= 0)
        return lenA < 0 || lenB < 0 ? fail(PMLG_ERR_ARG, "negative length") : PMLG_OK;
    return pmlg_polmat_mulmid(ctx, C, A, B, m, k, n, lenA, lenB, 0, lenA + lenB - 1, method);
}

void pmlg_profile_enable(pmlg_ctx ctx, int on)
{
    if (ctx)
        ctx->profile = on != 0;
}

pmlg_status pmlg_profile_read(pmlg_ctx ctx, double *ms, int64_t *launches)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    for (auto &pr : ctx->pending) {
        float t = 0.f;
        if (hipEventSynchronize(pr.b) == hipSuccess &&
            hipEventElapsedTime(&t, pr.a, pr.b) == hipSuccess) {
            ctx->ms[pr.stage] += t;
            ctx->launches[pr.stage] += 1;
        }
        (void)hipEventDestroy(pr.a);
        (void)hipEventDestroy(pr.b);
    }
    ctx->pending.clear();
    for (int i = 0; i < 3; ++i) {
        if (ms)
            ms[i] = ctx->ms[i];
        if (launches)
            launches[i] = ctx->launches[i];
        ctx->ms[i] = 0;
        ctx->launches[i] = 0;
    }
    return PMLG_OK;
}

pmlg_status pmlg_dev_copy(pmlg_ctx ctx, void *dst, const void *src, uint64_t bytes, void *stream)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16)
        return fail(PMLG_ERR_ARG, "pmlg_dev_copy needs 16-byte aligned pointers and size");
    DeviceGuard g(ctx->device);
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    const int64_t n16 = (int64_t)(bytes / 16);
    int64_t blocks = (n16 + 255) / 256;
    if (blocks > 16384)
        blocks = 16384;
    if (n16)
        hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                           static_cast<uint4 *>(dst), static_cast<const uint4 *>(src), n16);
    HIP_TRY(hipGetLastError(), "copy16_kernel");
    return PMLG_OK;
}

pmlg_status pmlg_dev_alloc(pmlg_ctx ctx, void **ptr, uint64_t bytes)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    DeviceGuard g(ctx->device);
    HIP_TRY(hipMalloc(ptr, bytes ? bytes : 8), "hipMalloc");
    return PMLG_OK;
}

void pmlg_dev_free(pmlg_ctx ctx, void *ptr)
{
    if (!ctx || !ptr)
        return;
    DeviceGuard g(ctx->device);
    (void)hipFree(ptr);
}

pmlg_status pmlg_memcpy_h2d(pmlg_ctx ctx, void *dst, const void *src, uint64_t bytes)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    DeviceGuard g(ctx->device);
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    return PMLG_OK;
}

pmlg_status pmlg_memcpy_d2h(pmlg_ctx ctx, void *dst, const void *src, uint64_t bytes)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    DeviceGuard g(ctx->device);
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return PMLG_OK;
}

}  // extern "C"
