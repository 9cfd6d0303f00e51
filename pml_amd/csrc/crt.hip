// Products modulo any p in [2, 2^64) whose roots of unity do not reach the
// transform length (non-FFT primes, composite moduli, every p >= 2^62): the
// integer product of
// the residues is computed modulo three FFT primes q1 > q2 > q3 in (2^59, 2^60)
// (below 2^60, so that their transforms run in the lazy butterfly mode) with
// the ordinary device path, recombined by Garner's mixed-radix CRT and
// reduced mod p.  A coefficient of A*B is at most k min(lenA, lenB) (p-1)^2 <
// 2^128 k min(lenA, lenB), below q1 q2 q3 > 2^179.99 for k min(lenA, lenB) < 2^51,
// so the recombined integer is the exact coefficient.  This is the
// multi-prime scheme of ntl-extras (mat_lzz_pX_multiply_3_primes.cpp:47-200),
// here on the GPU transforms.
#include <string>

#include "ctx.hpp"
#include "host_arith.hpp"

namespace pmlg {

namespace {

// the three largest primes c 2^32 + 1 below 2^60
constexpr uint64_t kCrtPrimes[3] = {1152921092289986561ull,   // 0x0fffffa000000001, 2^37 | q-1
                                    1152920989210771457ull,   // 0x0fffff8800000001, 2^35 | q-1
                                    1152920933376196609ull};  // 0x0fffff7b00000001, 2^32 | q-1

struct CrtConst {
    ModConst m2, m3, mp;     // moduli q2, q3 and p
    uint64_t q2, q3, p;
    uint64_t i12, i12q;      // q1^-1 mod q2 (Shoup companion for q2)
    uint64_t i123, i123q;    // (q1 q2)^-1 mod q3
    uint64_t q1_3, q1_3q;    // q1 mod q3
    uint64_t q1_p, q12_p;    // q1 mod p, q1 q2 mod p
    int big;                 // p >= 2^62: reductions for any 64-bit p
};

__device__ __forceinline__ uint64_t mulmod_p(uint64_t a, uint64_t b, const ModConst &mc)
{
    uint64_t hi, lo;
    mul128(a, b, hi, lo);
    return reduce128(hi, lo, mc);
}

// out[e*out_stride + d] for d < width from the residues c_i[e*width + d] mod q_i
__global__ void __launch_bounds__(256)
crt_kernel(uint64_t *__restrict__ out, int64_t out_stride, const uint64_t *__restrict__ c1,
           const uint64_t *__restrict__ c2, const uint64_t *__restrict__ c3, int64_t width,
           int64_t count, CrtConst k)
{
    const int64_t total = count * width;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = idx / width, d = idx - e * width;
        const uint64_t v1 = c1[idx];  // < q1 < 2 q2, 2 q3
        // v2 = (c2 - v1) / q1 mod q2
        uint64_t a = v1 >= k.q2 ? v1 - k.q2 : v1;
        uint64_t t = c2[idx] + (a > c2[idx] ? k.q2 : 0) - a;  // in [0, q2)
        const uint64_t v2 = shoup<uint64_t>(t, k.i12, k.i12q, k.q2);
        // v3 = (c3 - v1 - q1 v2) / (q1 q2) mod q3
        const uint64_t v1_3 = v1 >= k.q3 ? v1 - k.q3 : v1;
        const uint64_t v2_3 = v2 >= k.q3 ? v2 - k.q3 : v2;  // v2 < q2 < 2 q3
        uint64_t u = v1_3 + shoup<uint64_t>(v2_3, k.q1_3, k.q1_3q, k.q3);
        u = u >= k.q3 ? u - k.q3 : u;
        t = c3[idx] + (u > c3[idx] ? k.q3 : 0) - u;
        const uint64_t v3 = shoup<uint64_t>(t, k.i123, k.i123q, k.q3);
        // x = v1 + q1 v2 + q1 q2 v3, reduced mod p
        uint64_t x;
        if (k.big) {
            // p >= 2^62: v1, v2, v3 < 2^60 are already below p; sums may carry
            x = addmod_any(v1, mulmod_any(k.q1_p, v2, k.mp), k.p);
            x = addmod_any(x, mulmod_any(k.q12_p, v3, k.mp), k.p);
        } else {
            x = barrett(v1, k.mp.bm, k.p);
            x += mulmod_p(k.q1_p, v2, k.mp);
            x = x >= k.p ? x - k.p : x;
            x += mulmod_p(k.q12_p, v3, k.mp);
            x = x >= k.p ? x - k.p : x;
        }
        out[e * out_stride + d] = x;
    }
}

// dst[e*stride + i] = src[e*stride + i] mod q for i < len
__global__ void __launch_bounds__(256)
reduce_copy_kernel(uint64_t *__restrict__ dst, const uint64_t *__restrict__ src, int64_t count,
                   int64_t stride, int64_t len, uint64_t q, uint64_t bm)
{
    const int64_t total = count * len;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = idx / len, i = idx - e * len;
        dst[e * stride + i] = barrett(src[e * stride + i], bm, q);
    }
}

unsigned crt_grid(int64_t items)
{
    int64_t b = (items + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

}  // namespace

pmlg_status crt_product(pmlg_ctx ctx, uint64_t *dOut, int64_t out_stride, const uint64_t *dA,
                        int64_t strideA, const uint64_t *dB, int64_t strideB, int64_t m, int64_t k,
                        int64_t n, int64_t la, int64_t lb, int64_t nlo, int64_t hi, int method,
                        hipStream_t s)
{
    const int64_t width = hi - nlo;
    if (width <= 0 || m == 0 || n == 0)
        return PMLG_OK;
    for (int i = 0; i < 3; ++i)
        if (!ctx->crt[i]) {
            pmlg_status st = pmlg_ctx_create(&ctx->crt[i], kCrtPrimes[i], ctx->device);
            if (st != PMLG_OK)
                return st;
        }
    if ((unsigned __int128)k * (uint64_t)(la < lb ? la : lb) >= ((unsigned __int128)1 << 51))
        return fail(PMLG_ERR_UNSUPPORTED, "multi-prime path: k * min(lenA, lenB) must stay below "
                                          "2^51 for the three-prime recombination to be exact");
    const int64_t lc = la + lb - 1;
    const int64_t need = hi > lc - nlo ? hi : lc - nlo;
    if (ceil_log2(need) > ctx->crt[0]->max_log)
        return fail(PMLG_ERR_UNSUPPORTED, "transform length 2^" + std::to_string(ceil_log2(need)) +
                                              " exceeds the multi-prime path's limit 2^" +
                                              std::to_string(ctx->crt[0]->max_log));
    // residues below every q_i (p < 2^60) are valid inputs as they are;
    // otherwise each prime gets a reduced copy of A and B
    const bool reduce = ctx->p > kCrtPrimes[2];
    const size_t outw = (size_t)(m * n * width);
    HIP_TRY(ctx->crt_out.ensure(3 * outw * 8), "hipMalloc(crt residues)");
    uint64_t *cres = static_cast<uint64_t *>(ctx->crt_out.ptr);
    // one reduced copy of A and B per prime (the products run concurrently)
    const size_t per = (size_t)(m * k * strideA + k * n * strideB);
    if (reduce)
        HIP_TRY(ctx->crt_in.ensure(3 * per * 8 + 256), "hipMalloc(crt inputs)");
    // The three products are independent: they run on the sub-contexts'
    // streams, forked from s by an event and joined back before the
    // recombination, so each chain's launch gaps are filled by the others.
    // A stream being captured keeps them in sequence on s.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(s, &cs), "hipStreamIsCapturing");
    const bool fork = cs == hipStreamCaptureStatusNone;
    if (fork) {
        if (!ctx->crt_fork) {
            HIP_TRY(hipEventCreateWithFlags(&ctx->crt_fork, hipEventDisableTiming), "hipEventCreate");
            for (hipEvent_t &e : ctx->crt_join)
                HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
        }
        HIP_TRY(hipEventRecord(ctx->crt_fork, s), "hipEventRecord");
    }
    // the modulus of each product is q_i: the MFMA kernel is used where its
    // bound allows it, as for any other prime
    const int sub_method = method == PMLG_METHOD_PLAIN ? PMLG_METHOD_PLAIN : PMLG_METHOD_AUTO;
    for (int i = 0; i < 3; ++i) {
        pmlg_ctx q = ctx->crt[i];
        hipStream_t si = fork ? q->stream : s;
        if (fork)
            HIP_TRY(hipStreamWaitEvent(si, ctx->crt_fork, 0), "hipStreamWaitEvent");
        const uint64_t *a = dA, *b = dB;
        if (reduce) {
            uint64_t *rA = static_cast<uint64_t *>(ctx->crt_in.ptr) + i * per;
            uint64_t *rB = rA + m * k * strideA;
            hipLaunchKernelGGL(reduce_copy_kernel, dim3(crt_grid(m * k * la)), dim3(256), 0, si, rA, dA,
                               m * k, strideA, la, q->p, q->mc.bm);
            hipLaunchKernelGGL(reduce_copy_kernel, dim3(crt_grid(k * n * lb)), dim3(256), 0, si, rB, dB,
                               k * n, strideB, lb, q->p, q->mc.bm);
            HIP_TRY(hipGetLastError(), "reduce_copy_kernel");
            a = rA;
            b = rB;
        }
        pmlg_status st = mulmid_dev_strided(q, cres + i * outw, width, a, strideA, b, strideB, m, k,
                                            n, la, lb, nlo, hi, sub_method, si);
        if (st != PMLG_OK)
            return st;
        if (fork)
            HIP_TRY(hipEventRecord(ctx->crt_join[i], si), "hipEventRecord");
    }
    if (fork)
        for (hipEvent_t e : ctx->crt_join)
            HIP_TRY(hipStreamWaitEvent(s, e, 0), "hipStreamWaitEvent");
    // Garner constants
    const uint64_t q1 = kCrtPrimes[0], q2 = kCrtPrimes[1], q3 = kCrtPrimes[2], p = ctx->p;
    CrtConst c{};
    c.m2 = ctx->crt[1]->mc;
    c.m3 = ctx->crt[2]->mc;
    c.mp = ctx->mc;
    c.q2 = q2;
    c.q3 = q3;
    c.p = p;
    c.i12 = host_invmod(q1 % q2, q2);
    c.i12q = host_shoup_pre64(c.i12, q2);
    c.i123 = host_invmod(host_mulmod(q1 % q3, q2 % q3, q3), q3);
    c.i123q = host_shoup_pre64(c.i123, q3);
    c.q1_3 = q1 % q3;
    c.q1_3q = host_shoup_pre64(c.q1_3, q3);
    c.q1_p = q1 % p;
    c.q12_p = host_mulmod(q1 % p, q2 % p, p);
    c.big = p >= ((uint64_t)1 << 62) ? 1 : 0;
    hipLaunchKernelGGL(crt_kernel, dim3(crt_grid((int64_t)outw)), dim3(256), 0, s, dOut, out_stride,
                       cres, cres + outw, cres + 2 * outw, width, m * n, c);
    HIP_TRY(hipGetLastError(), "crt_kernel");
    return PMLG_OK;
}

}  // namespace pmlg
