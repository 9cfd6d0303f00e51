// Host <-> device transfers of polynomial matrices held entry by entry in
// host memory (FLINT's layout: one coefficient array per entry), without an
// intermediate packed host copy.  Entries are packed straight into pinned
// staging chunks and copied to the device's entry-major layout by DMA, while
// the next chunk is being packed; results come back the same way and are
// written into the caller's entries with their normalised lengths.
//
// stage_threads() host threads each own two pinned chunks and one copy
// stream, and take every stage_threads()-th chunk, so packing (a host memory
// copy) runs on several cores while the DMA engines stay busy.  Two threads
// with 2 MiB chunks measured best for 32x32 matrices of length 4096 (more
// threads contend for the copy engines); PMLG_STAGE_THREADS and
// PMLG_STAGE_MIB override them.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "ctx.hpp"

namespace pmlg {

static int env_int(const char *name, int dflt, int lo, int hi)
{
    const char *v = getenv(name);
    if (!v || !*v)
        return dflt;
    const int x = atoi(v);
    return x < lo ? lo : (x > hi ? hi : x);
}

int stage_threads()
{
    static const int t = env_int("PMLG_STAGE_THREADS", 2, 1, kMaxStageThreads);
    return t;
}

size_t stage_bytes()
{
    static const size_t b = (size_t)env_int("PMLG_STAGE_MIB", 2, 1, 256) << 20;
    return b;
}

namespace {

// make sure every stager exists and has chunks of at least `bytes`
pmlg_status stagers_ready(pmlg_ctx ctx, size_t bytes)
{
    if (ctx->stagers.empty())
        ctx->stagers.resize((size_t)stage_threads());
    for (Stager &S : ctx->stagers) {
        if (!S.s) {
            HIP_TRY(hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking), "hipStreamCreate(stager)");
            for (int b = 0; b < 2; ++b)
                HIP_TRY(hipEventCreateWithFlags(&S.ev[b], hipEventDisableTiming), "hipEventCreate");
        }
        if (S.bytes < bytes) {
            for (int b = 0; b < 2; ++b) {
                if (S.pin[b]) {
                    (void)hipEventSynchronize(S.ev[b]);
                    (void)hipHostFree(S.pin[b]);
                    S.pin[b] = nullptr;
                }
            }
            S.bytes = 0;
            for (int b = 0; b < 2; ++b)
                HIP_TRY(hipHostMalloc(&S.pin[b], bytes, hipHostMallocDefault), "hipHostMalloc(stager)");
            S.bytes = bytes;
            S.used[0] = S.used[1] = false;
        }
    }
    return PMLG_OK;
}

size_t chunk_bytes(int64_t width)
{
    return std::max<size_t>(stage_bytes(), (size_t)width * 8);
}

// runs fn(t) on `active` threads (the calling thread takes t = 0) and
// returns the first failure
template <typename Fn>
pmlg_status run_threads(int active, Fn fn)
{
    std::vector<pmlg_status> st((size_t)active, PMLG_OK);
    std::vector<std::thread> th;
    for (int t = 1; t < active; ++t)
        th.emplace_back([&, t] { st[(size_t)t] = fn(t); });
    st[0] = fn(0);
    for (auto &x : th)
        x.join();
    for (pmlg_status s : st)
        if (s != PMLG_OK)
            return s;
    return PMLG_OK;
}

}  // namespace

void stagers_release(pmlg_ctx ctx)
{
    for (Stager &S : ctx->stagers) {
        if (S.s)
            (void)hipStreamSynchronize(S.s);
        for (int b = 0; b < 2; ++b) {
            if (S.pin[b])
                (void)hipHostFree(S.pin[b]);
            if (S.ev[b])
                (void)hipEventDestroy(S.ev[b]);
        }
        if (S.s)
            (void)hipStreamDestroy(S.s);
    }
    ctx->stagers.clear();
}

// dst (device) entry e at dst + e*width <- the first min(len[e], width) words
// of ptr[e], zero padded.  Returns when the copies are complete.
pmlg_status stage_upload(pmlg_ctx ctx, uint64_t *dst, const uint64_t *const *ptr,
                         const int64_t *len, int64_t count, int64_t width)
{
    if (count <= 0 || width <= 0)
        return PMLG_OK;
    const size_t cb = chunk_bytes(width);
    pmlg_status st = stagers_ready(ctx, cb);
    if (st != PMLG_OK)
        return st;
    const int64_t per = (int64_t)(cb / ((size_t)width * 8));
    const int64_t chunks = (count + per - 1) / per;
    const int active = (int)std::min<int64_t>((int64_t)ctx->stagers.size(), chunks);
    const int device = ctx->device;
    return run_threads(active, [&](int t) -> pmlg_status {
        DeviceGuard g(device);
        Stager &S = ctx->stagers[(size_t)t];
        int64_t local = 0;
        for (int64_t c = t; c < chunks; c += active, ++local) {
            const int b = (int)(local & 1);
            if (S.used[b])
                HIP_TRY(hipEventSynchronize(S.ev[b]), "hipEventSynchronize(stager)");
            const int64_t e0 = c * per, e1 = std::min(count, e0 + per);
            uint64_t *pin = static_cast<uint64_t *>(S.pin[b]);
            for (int64_t e = e0; e < e1; ++e) {
                uint64_t *d = pin + (e - e0) * width;
                const int64_t l = std::max<int64_t>(0, std::min(len[e], width));
                if (l)
                    std::memcpy(d, ptr[e], (size_t)l * 8);
                if (l < width)
                    std::memset(d + l, 0, (size_t)(width - l) * 8);
            }
            HIP_TRY(hipMemcpyAsync(dst + e0 * width, pin, (size_t)((e1 - e0) * width) * 8,
                                   hipMemcpyHostToDevice, S.s),
                    "H2D(stager)");
            HIP_TRY(hipEventRecord(S.ev[b], S.s), "hipEventRecord(stager)");
            S.used[b] = true;
        }
        HIP_TRY(hipStreamSynchronize(S.s), "hipStreamSynchronize(stager)");
        return PMLG_OK;
    });
}

// ptr[e] (room for width words) <- entry e of src (device, at src + e*width),
// normalised: lenOut[e] is its length without trailing zeros and only those
// words are written.  The copies wait for work already enqueued on `after`.
pmlg_status stage_download(pmlg_ctx ctx, uint64_t *const *ptr, int64_t *lenOut,
                           const uint64_t *src, int64_t count, int64_t width, hipStream_t after,
                           int64_t zero_to)
{
    if (count <= 0)
        return PMLG_OK;
    if (width <= 0) {
        for (int64_t e = 0; e < count; ++e) {
            lenOut[e] = 0;
            if (zero_to > 0)
                std::memset(ptr[e], 0, (size_t)zero_to * 8);
        }
        return PMLG_OK;
    }
    const size_t cb = chunk_bytes(width);
    pmlg_status st = stagers_ready(ctx, cb);
    if (st != PMLG_OK)
        return st;
    hipEvent_t ready;
    HIP_TRY(hipEventCreateWithFlags(&ready, hipEventDisableTiming), "hipEventCreate");
    hipError_t er = hipEventRecord(ready, after);
    if (er != hipSuccess) {
        (void)hipEventDestroy(ready);
        return hip_fail(er, "hipEventRecord");
    }
    const int64_t per = (int64_t)(cb / ((size_t)width * 8));
    const int64_t chunks = (count + per - 1) / per;
    const int active = (int)std::min<int64_t>((int64_t)ctx->stagers.size(), chunks);
    const int device = ctx->device;
    st = run_threads(active, [&](int t) -> pmlg_status {
        DeviceGuard g(device);
        Stager &S = ctx->stagers[(size_t)t];
        HIP_TRY(hipStreamWaitEvent(S.s, ready, 0), "hipStreamWaitEvent");
        auto unpack = [&](int64_t c, int b) -> pmlg_status {
            HIP_TRY(hipEventSynchronize(S.ev[b]), "hipEventSynchronize(stager)");
            const int64_t e0 = c * per, e1 = std::min(count, e0 + per);
            const uint64_t *pin = static_cast<const uint64_t *>(S.pin[b]);
            for (int64_t e = e0; e < e1; ++e) {
                const uint64_t *s = pin + (e - e0) * width;
                int64_t l = width;
                while (l > 0 && s[l - 1] == 0)
                    --l;
                if (l)
                    std::memcpy(ptr[e], s, (size_t)l * 8);
                if (zero_to > l)
                    std::memset(ptr[e] + l, 0, (size_t)(zero_to - l) * 8);
                lenOut[e] = l;
            }
            return PMLG_OK;
        };
        int64_t prev = -1, local = 0;
        int pb = 0;
        for (int64_t c = t; c < chunks; c += active, ++local) {
            const int b = (int)(local & 1);
            const int64_t e0 = c * per, e1 = std::min(count, e0 + per);
            HIP_TRY(hipMemcpyAsync(S.pin[b], src + e0 * width, (size_t)((e1 - e0) * width) * 8,
                                   hipMemcpyDeviceToHost, S.s),
                    "D2H(stager)");
            HIP_TRY(hipEventRecord(S.ev[b], S.s), "hipEventRecord(stager)");
            S.used[b] = true;
            if (prev >= 0) {
                pmlg_status s2 = unpack(prev, pb);
                if (s2 != PMLG_OK)
                    return s2;
            }
            prev = c;
            pb = b;
        }
        if (prev >= 0)
            return unpack(prev, pb);
        return PMLG_OK;
    });
    (void)hipEventDestroy(ready);
    return st;
}

}  // namespace pmlg

using namespace pmlg;

extern "C" {

pmlg_status pmlg_polmat_mulmid_entries(pmlg_ctx ctx, uint64_t *const *out, int64_t *lenOut,
                                       const uint64_t *const *A, const int64_t *lenA,
                                       const uint64_t *const *B, const int64_t *lenB, int64_t m,
                                       int64_t k, int64_t n, int64_t nlo, int64_t nhi, int method)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    if (m < 0 || k < 0 || n < 0 || nlo < 0)
        return fail(PMLG_ERR_ARG, "negative dimension or nlo");
    if ((m * n > 0 && (!out || !lenOut)) || (m * k > 0 && (!A || !lenA)) ||
        (k * n > 0 && (!B || !lenB)))
        return fail(PMLG_ERR_ARG, "null entry array");
    const int64_t width = nhi - nlo;
    if (m == 0 || n == 0)
        return PMLG_OK;
    // only coefficients below nhi of each operand reach [nlo, nhi)
    int64_t la = 0, lb = 0;
    for (int64_t e = 0; e < m * k; ++e)
        la = std::max(la, lenA[e]);
    for (int64_t e = 0; e < k * n; ++e)
        lb = std::max(lb, lenB[e]);
    la = std::min(la, std::max<int64_t>(nhi, 0));
    lb = std::min(lb, std::max<int64_t>(nhi, 0));
    if (width <= 0 || k == 0 || la == 0 || lb == 0 || nlo >= la + lb - 1) {
        for (int64_t e = 0; e < m * n; ++e)
            lenOut[e] = 0;
        return PMLG_OK;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    hipStream_t s = ctx->stream;
    const size_t ba = (size_t)(m * k * la) * 8, bb = (size_t)(k * n * lb) * 8;
    const size_t bc = (size_t)(m * n * width) * 8;
    HIP_TRY(ctx->in_a.ensure(ba), "hipMalloc(A)");
    HIP_TRY(ctx->in_b.ensure(bb), "hipMalloc(B)");
    HIP_TRY(ctx->out_c.ensure(bc), "hipMalloc(C)");
    // the buffers may still be read by work enqueued earlier on the context's stream
    HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
    st = stage_upload(ctx, static_cast<uint64_t *>(ctx->in_a.ptr), A, lenA, m * k, la);
    if (st != PMLG_OK)
        return st;
    st = stage_upload(ctx, static_cast<uint64_t *>(ctx->in_b.ptr), B, lenB, k * n, lb);
    if (st != PMLG_OK)
        return st;
    st = mulmid_dev_locked(ctx, static_cast<uint64_t *>(ctx->out_c.ptr),
                           static_cast<const uint64_t *>(ctx->in_a.ptr),
                           static_cast<const uint64_t *>(ctx->in_b.ptr), m, k, n, la, lb, nlo, nhi,
                           method, s);
    if (st != PMLG_OK) {
        (void)hipStreamSynchronize(s);
        return st;
    }
    return stage_download(ctx, out, lenOut, static_cast<const uint64_t *>(ctx->out_c.ptr), m * n,
                          width, s);
}

pmlg_status pmlg_polmat_mul_entries(pmlg_ctx ctx, uint64_t *const *C, int64_t *lenC,
                                    const uint64_t *const *A, const int64_t *lenA,
                                    const uint64_t *const *B, const int64_t *lenB, int64_t m,
                                    int64_t k, int64_t n, int method)
{
    int64_t la = 0, lb = 0;
    for (int64_t e = 0; A && lenA && e < m * k; ++e)
        la = std::max(la, lenA[e]);
    for (int64_t e = 0; B && lenB && e < k * n; ++e)
        lb = std::max(lb, lenB[e]);
    const int64_t hi = (la > 0 && lb > 0) ? la + lb - 1 : 0;
    return pmlg_polmat_mulmid_entries(ctx, C, lenC, A, lenA, B, lenB, m, k, n, 0, hi, method);
}

}  // extern "C"
