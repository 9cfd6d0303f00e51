// Batched number-theoretic transforms over Z/pZ at 2^k-th roots of unity.
//
// Forward: decimation in frequency (natural order in, bit-reversed order out
// inside each LDS sub-transform).  Inverse: decimation in time, exact inverse
// of the forward up to the factor N.  Pointwise products work position by
// position, so the evaluation order never has to be undone.
//
// One workgroup holds `nb` sub-transforms of size M = 2^logM in LDS and runs
// every radix-2 stage there.  When N = R*M exceeds the LDS limit, the first
// radix-R stage is applied while loading (y_g[i] = w^(i g) sum_t x[i+Mt] w_R^(tg),
// whose size-M transform is X[R k + g]); the inverse runs the R size-M inverse
// transforms and a merge pass x[o] = N^-1 sum_g w^(-o g) z_g[o mod M].
#include "internal.hpp"

namespace pmlg {

namespace {

constexpr int kThreads = 256;
#ifndef PMLG_INV_WAVES
#define PMLG_INV_WAVES 4
#endif
constexpr int kInvWaves = PMLG_INV_WAVES;  // minimum waves per SIMD of the inverse register kernels
// minimum waves per SIMD of the radix-32 forward register kernels: their
// 256-thread workgroups alone ask for one, which lets the register allocator
// drift past 256 registers (one wave per SIMD, 1.6x slower) on small changes
// of the surrounding code; two keep them at or below 256.
constexpr int kFwdWaves32 = 2;

// Point-blocked evaluation layout: position `pos` of entry `e` (of `count`
// entries) lives at ((pos >> logP) * count + e) * P + (pos mod P), P = 2^logP,
// so that the P values of every entry at P consecutive points form one run
// and all entries of one point block are contiguous.
__device__ __forceinline__ int64_t blk_addr(int64_t e, int64_t pos, int64_t count, int logP)
{
    return (((pos >> logP) * count + e) << logP) | (pos & ((1 << logP) - 1));
}

// Workgroups are dispatched round-robin over the 8 XCDs; give the workgroups
// of one XCD consecutive units, so that neighbouring entries (which share
// lines when the runs are shorter than a line) meet in the same L2.
__device__ __forceinline__ int64_t xcd_unit(int64_t b, int64_t units)
{
    return (units & 7) ? b : (b & 7) * (units >> 3) + (b >> 3);
}

template <typename W>
__device__ __forceinline__ W tw_mul(W x, const W *w, const W *wq, int64_t idx, W p)
{
    return shoup_lazy(x, w[idx], wq[idx], p);
}

// The last two inverse stages of a 3/4-truncated product.  The forward's
// first two stages map the coefficients x0 = c[i], x1 = c[i + N/4], x2 =
// c[i + N/2], x3 = c[i + 3N/4] (i < N/4) to Y0 = x0 + x1 + x2 + x3, Y1 = (x0 +
// x2 - x1 - x3) w^(2i), Y2 = w^i ((x0 - x2) + s (x1 - x3)), Y3 (dropped),
// s = w^(N/4).  A product of length <= 3N/4 has x3 = 0, so with the inverse's
// earlier stages leaving (N/4) Y0, (N/4) Y1, (N/4) Y2, the coefficients are
//   x1 = 2/N (Y0 - u),  x0 = S + D - (s/2) x1,  x2 = S - D + (s/2) x1,
// u = w^(-2i) Y1, S = (Y0 + u)/N, D = 2/N w^(-i) Y2, scaled like the other
// outputs (N^-1, or N^-1 2^64 after a REDC pointwise product).  Any word is
// a valid input (each is first taken through a Shoup product); the outputs
// are in [0, p).
template <typename W>
__device__ __forceinline__ void trunc_solve(W &y0, W &y1, W &y2, int64_t i, const NttTab<W> &tab,
                                            W p)
{
    const W a = reduce_2p<W>(shoup_lazy(y0, tab.itw[0], tab.itwq[0], p), p);
    const W u = reduce_2p<W>(shoup_lazy(y1, tab.itw[2 * i], tab.itwq[2 * i], p), p);
    const W v = reduce_2p<W>(shoup_lazy(y2, tab.itw[i], tab.itwq[i], p), p);
    const W d = a >= u ? a - u : a + (p - u);
    W su = a + u;
    su = su >= p ? su - p : su;
    const W x1 = shoup<W>(d, tab.ninv2, tab.ninv2q, p);
    const W S = shoup<W>(su, tab.ninv, tab.ninvq, p);
    const W D = shoup<W>(v, tab.ninv2, tab.ninv2q, p);
    const W h = shoup<W>(x1, tab.thalf, tab.thalfq, p);
    W sd = S + D;
    sd = sd >= p ? sd - p : sd;
    W md = S >= D ? S - D : S + (p - D);
    y0 = sd >= h ? sd - h : sd + (p - h);
    md = md + h;
    y2 = md >= p ? md - p : md;
    y1 = x1;
}

// Output correction of a truncated product inverted as a whole: with any
// values at the dropped evaluations (frequencies 3 mod 4, contributing U3'
// where U3 was due), the scaled outputs are c_t' = c_t + s^(-3t) (U3' - U3)/4
// for t = 0..3 (c_t = c[i + tN/4], s = w^(N/4), s^2 = -1), and c_3 = 0 makes
// (U3' - U3)/4 = s c_3'.  So c_0 = c_0' - s c_3', c_1 = c_1' + c_3', c_2 =
// c_2' + s c_3', from scaled outputs in [0, p).
template <typename W>
__device__ __forceinline__ void trunc_correct(W &c0, W &c1, W &c2, W c3, const NttTab<W> &tab, W p)
{
    const W sc = shoup<W>(c3, tab.tquart, tab.tquartq, p);
    c0 = c0 >= sc ? c0 - sc : c0 + (p - sc);
    c1 = c1 + c3;
    c1 = c1 >= p ? c1 - p : c1;
    c2 = c2 + sc;
    c2 = c2 >= p ? c2 - p : c2;
}

// Units u0 .. u0+nb-1 (of `units` = count << logR) of one operand: the
// workgroup body shared by the direct and the segment-table launches.
template <typename Win, typename W>
__device__ __forceinline__ void fwd_lds_body(const Win *__restrict__ in, int64_t in_stride,
                                             int64_t in_len, W *__restrict__ out, int64_t units,
                                             int64_t u0, int logN, int logR, int lognb, int logP,
                                             const NttTab<W> &tab, W p, bool trunc = false)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    W *s = reinterpret_cast<W *>(smem_raw);
    const int logM = logN - logR;
    const int M = 1 << logM;
    const int R = 1 << logR;
    const int nb = 1 << lognb;
    const int64_t N = (int64_t)1 << logN;
    const W p2 = p + p;

    // load (with the radix-R split when R > 1)
    for (int idx = threadIdx.x; idx < nb * M; idx += kThreads) {
        const int t = idx >> logM;
        const int i1 = idx & (M - 1);
        const int64_t u = u0 + t;
        W v = 0;
        if (u < units) {
            const int64_t e = u >> logR;
            const int g = (int)(u & (R - 1));
            const Win *x = in + e * in_stride;
            if (R == 1) {
                v = (i1 < in_len) ? (W)x[i1] : (W)0;
            } else {
                W acc = 0;
                for (int t2 = 0; t2 < R; ++t2) {
                    const int64_t src = (int64_t)i1 + (int64_t)M * t2;
                    if (src < in_len) {
                        const int64_t widx = ((int64_t)M * g * t2) & (N - 1);
                        W xv = (W)x[src];
                        acc = add_2p<W>(acc, tw_mul<W>(xv, tab.tw, tab.twq, widx, p), p2);
                    }
                }
                v = tw_mul<W>(acc, tab.tw, tab.twq, (int64_t)i1 * g, p);
            }
        }
        s[idx] = v;
    }
    __syncthreads();

    // DIF radix-2 stages; sub-transform root is w^R, twiddle w_M^(j M/2h) = w^(j << (logN-lh-1))
    const int halfM = M >> 1;
    for (int lh = logM - 1; lh >= 0; --lh) {
        const int h = 1 << lh;
        const int sh = logN - lh - 1;
        for (int b = threadIdx.x; b < nb * halfM; b += kThreads) {
            const int t = b >> (logM - 1);
            const int bb = b & (halfM - 1);
            const int j = bb & (h - 1);
            const int i = t * M + ((bb >> lh) << (lh + 1)) + j;
            const W a = s[i], c = s[i + h];
            s[i] = add_2p<W>(a, c, p2);
            s[i + h] = tw_mul<W>(sub_2p<W>(a, c, p2), tab.tw, tab.twq, (int64_t)j << sh, p);
        }
        __syncthreads();
    }

    // store: unit u = (e, g) -> out[e*N + g*M + i]; a 3/4-truncated transform
    // (R = 1) keeps the positions below 3M/4
    for (int idx = threadIdx.x; idx < nb * M; idx += kThreads) {
        const int t = idx >> logM;
        const int i1 = idx & (M - 1);
        const int64_t u = u0 + t;
        if (u < units && !(trunc && R == 1 && i1 >= 3 * (M >> 2))) {
            const int64_t e = u >> logR;
            const int g = (int)(u & (R - 1));
            out[blk_addr(e, (int64_t)g * M + i1, units >> logR, logP)] = reduce_2p<W>(s[idx], p);
        }
    }
}

template <typename Win, typename W>
__global__ void __launch_bounds__(kThreads)
ntt_fwd_kernel(const Win *__restrict__ in, int64_t in_stride, int64_t in_len,
               W *__restrict__ out, int64_t units, int logN, int logR, int lognb, int logP,
               NttTab<W> tab, W p, int trunc)
{
    fwd_lds_body<Win, W>(in, in_stride, in_len, out, units, (int64_t)blockIdx.x << lognb, logN,
                         logR, lognb, logP, tab, p, trunc != 0);
}

template <typename W, typename Wout>
__device__ __forceinline__ void inv_lds_body(const W *__restrict__ in, Wout *__restrict__ out,
                                             int64_t out_stride, int64_t lo, int64_t hi,
                                             W *__restrict__ tmp, int64_t units, int64_t u0,
                                             int logN, int logR, int lognb, int logP,
                                             const NttTab<W> &tab, W p, bool trunc = false)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    W *s = reinterpret_cast<W *>(smem_raw);
    const int logM = logN - logR;
    const int M = 1 << logM;
    const int R = 1 << logR;
    const int nb = 1 << lognb;
    const int64_t N = (int64_t)1 << logN;
    const W p2 = p + p;

    for (int idx = threadIdx.x; idx < nb * M; idx += kThreads) {
        const int t = idx >> logM;
        const int i1 = idx & (M - 1);
        const int64_t u = u0 + t;
        W v = 0;
        if (u < units) {
            const int64_t e = u >> logR;
            const int g = (int)(u & (R - 1));
            if (!(trunc && i1 >= 3 * (M >> 2))) {  // a truncated product's dropped quarter
                v = in[blk_addr(e, (int64_t)g * M + i1, units >> logR, logP)];
                v = v >= p2 ? v - p2 : v;  // products may arrive in [0, 4p)
            }
        }
        s[idx] = v;
    }
    __syncthreads();

    // DIT radix-2 stages with inverse twiddles, h = 1 .. M/2
    const int halfM = M >> 1;
    for (int lh = 0; lh < logM; ++lh) {
        const int h = 1 << lh;
        const int sh = logN - lh - 1;
        for (int b = threadIdx.x; b < nb * halfM; b += kThreads) {
            const int t = b >> (logM - 1);
            const int bb = b & (halfM - 1);
            const int j = bb & (h - 1);
            const int i = t * M + ((bb >> lh) << (lh + 1)) + j;
            const W a = s[i];
            const W c = tw_mul<W>(s[i + h], tab.itw, tab.itwq, (int64_t)j << sh, p);
            s[i] = add_2p<W>(a, c, p2);
            W d = a + p2 - c;  // (0, 4p)
            s[i + h] = d >= p2 ? d - p2 : d;
        }
        __syncthreads();
    }

    if (trunc && R == 1) {
        // scaled outputs, corrected by the dropped quarter (trunc_correct)
        const int M4 = M >> 2;
        for (int b = threadIdx.x; b < nb * M4; b += kThreads) {
            const int t = b >> (logM - 2);
            const int i = b & (M4 - 1);
            W *q = s + t * M + i;
            W c0 = shoup<W>(q[0], tab.ninv, tab.ninvq, p), c1 = shoup<W>(q[M4], tab.ninv, tab.ninvq, p);
            W c2 = shoup<W>(q[2 * M4], tab.ninv, tab.ninvq, p);
            trunc_correct<W>(c0, c1, c2, shoup<W>(q[3 * M4], tab.ninv, tab.ninvq, p), tab, p);
            q[0] = c0;
            q[M4] = c1;
            q[2 * M4] = c2;
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < nb * M; idx += kThreads) {
        const int t = idx >> logM;
        const int i1 = idx & (M - 1);
        const int64_t u = u0 + t;
        if (u >= units)
            continue;
        const int64_t e = u >> logR;
        if (R == 1) {
            if (i1 >= lo && i1 < hi)
                out[e * out_stride + (i1 - lo)] =
                    (Wout)(trunc && R == 1 ? s[idx] : shoup<W>(s[idx], tab.ninv, tab.ninvq, p));
        } else {
            const int g = (int)(u & (R - 1));
            tmp[e * N + (int64_t)g * M + i1] = s[idx];
        }
    }
}

template <typename W, typename Wout>
__global__ void __launch_bounds__(kThreads)
ntt_inv_kernel(const W *__restrict__ in, Wout *__restrict__ out, int64_t out_stride, int64_t lo,
               int64_t hi, W *__restrict__ tmp, int64_t units, int logN, int logR, int lognb,
               int logP, NttTab<W> tab, W p, int trunc)
{
    inv_lds_body<W, Wout>(in, out, out_stride, lo, hi, tmp, units, (int64_t)blockIdx.x << lognb,
                          logN, logR, lognb, logP, tab, p, trunc != 0);
}

// x[o] = N^-1 * sum_g w^(-(o g) mod N) z_g[o mod M], o in [lo, hi)
template <typename W, typename Wout>
__device__ __forceinline__ void merge_one(const W *__restrict__ z, Wout *__restrict__ out,
                                          int64_t out_stride, int64_t lo, int64_t width,
                                          int64_t idx, int logN, int logR, const NttTab<W> &tab,
                                          W p)
{
    const int64_t N = (int64_t)1 << logN;
    const int64_t M = (int64_t)1 << (logN - logR);
    const int R = 1 << logR;
    const W p2 = p + p;
    const int64_t e = idx / width;
    const int64_t o = lo + (idx - e * width);
    const int64_t i1 = o & (M - 1);
    const W *ze = z + e * N;
    // z holds the row inverses' lazy words (any value below 2^bits(W)): the
    // first term too goes through a Shoup product (by w^0 = 1), into [0, 2p)
    W acc = tw_mul<W>(ze[i1], tab.itw, tab.itwq, 0, p);
    for (int g = 1; g < R; ++g) {
        const int64_t widx = (o * g) & (N - 1);
        acc = add_2p<W>(acc, tw_mul<W>(ze[(int64_t)g * M + i1], tab.itw, tab.itwq, widx, p), p2);
    }
    out[e * out_stride + (o - lo)] = (Wout)shoup<W>(acc, tab.ninv, tab.ninvq, p);
}

// Merge of a truncated product split in four (N = 4M, the sub-transform of
// frequencies 3 mod 4 dropped): item (e, i), i < M, turns z_0[i], z_1[i],
// z_2[i] (the sub-inverses, z_g = M w^(i g) sum_t c[i + tM] w^(tMg)) into
// c[i], c[i + M], c[i + 2M] (c[i + 3M] = 0): the relations of trunc_solve
// with z_2 in the role of the frequencies 2 mod 4 and z_1 of 1 mod 4.
template <typename W, typename Wout>
__device__ __forceinline__ void merge_trunc_one(const W *__restrict__ z, Wout *__restrict__ out,
                                                int64_t out_stride, int64_t lo, int64_t hi,
                                                int64_t idx, int logN, const NttTab<W> &tab, W p)
{
    const int64_t N = (int64_t)1 << logN, M = N >> 2;
    const int64_t e = idx / M, i = idx - e * M;
    const W *ze = z + e * N;
    W c0 = ze[i], c1 = ze[2 * M + i], c2 = ze[M + i];
    trunc_solve<W>(c0, c1, c2, i, tab, p);
    Wout *o = out + e * out_stride;
    if (i >= lo && i < hi)
        o[i - lo] = (Wout)c0;
    if (i + M >= lo && i + M < hi)
        o[i + M - lo] = (Wout)c1;
    if (i + 2 * M >= lo && i + 2 * M < hi)
        o[i + 2 * M - lo] = (Wout)c2;
}

template <typename W, typename Wout>
__global__ void __launch_bounds__(kThreads)
ntt_merge_trunc_kernel(const W *__restrict__ z, Wout *__restrict__ out, int64_t out_stride,
                       int64_t lo, int64_t hi, int64_t count, int logN, NttTab<W> tab, W p)
{
    const int64_t total = count << (logN - 2);
    for (int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * kThreads)
        merge_trunc_one<W, Wout>(z, out, out_stride, lo, hi, idx, logN, tab, p);
}

template <typename W, typename Wout>
__global__ void __launch_bounds__(kThreads)
ntt_merge_kernel(const W *__restrict__ z, Wout *__restrict__ out, int64_t out_stride, int64_t lo,
                 int64_t hi, int64_t count, int logN, int logR, NttTab<W> tab, W p)
{
    const int64_t width = hi - lo;
    const int64_t total = count * width;
    for (int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * kThreads)
        merge_one<W, Wout>(z, out, out_stride, lo, width, idx, logN, logR, tab, p);
}

inline int pick_lognb(int logM)
{
    const int want = 11;  // at least 2048 elements per workgroup
    return logM >= want ? 0 : want - logM;
}

// ===========================================================================
// Register-resident transforms for 2^11 <= M <= 2^14.
//
// Each of the T = M / E threads keeps E = 2^LOGE elements in registers.  The
// L = log2 M radix-2 stages are grouped into passes of LOGE stages (the last
// one possibly shorter); a pass starting at stage S0 with R stages acts on the
// index bits [BLO, BLO+R), BLO = L-S0-R, and every thread owns whole groups of
// 2^R elements that differ only in those bits.  Between passes the elements go
// through LDS, laid out with one padding word every 32 to spread the banks.
// Stage s pairs bit b = L-1-s; the twiddle of a pair whose lower index is i is
// w_M^((i mod 2^b) 2^s), read from the per-stage table at x = i mod 2^b.
// ===========================================================================

__device__ __forceinline__ int swz(int a)
{
    return a + (a >> 5);
}

template <typename W>
__device__ __forceinline__ void load_tw(const W *__restrict__ tab, int idx, W &w, W &wq)
{
    if constexpr (sizeof(W) == 8) {
        const uint4 v = *reinterpret_cast<const uint4 *>(tab + 2 * idx);
        w = ((uint64_t)v.y << 32) | v.x;
        wq = ((uint64_t)v.w << 32) | v.z;
    } else {
        const uint2 v = *reinterpret_cast<const uint2 *>(tab + 2 * idx);
        w = v.x;
        wq = v.y;
    }
}

// Butterfly constants.  Strict mode keeps values in [0, 2p) (any p < 2^62).
// Lazy mode (64-bit words, p < 2^60, so that 16p < 2^64) lets values grow up
// to 16p and tracks a bound for every register at compile time (the
// fwd_bound / inv_bound rules below): the Shoup product returns [0, 4p) for
// any 64-bit input, so only sums grow, and a conditional subtraction of 8p is
// placed only where a bound would pass 16p.
template <typename W>
struct Bfly {
    W p, p2, p4, p8;
    W np, np2, np4, np8;  // -p, -2p, -4p, -8p mod 2^bits(W)
    __device__ explicit Bfly(W pp)
        : p(pp), p2(pp + pp), p4(4 * pp), p8(8 * pp), np((W)0 - pp), np2((W)0 - 2 * pp),
          np4((W)0 - 4 * pp), np8((W)0 - 8 * pp)
    {
        // keep the negatives opaque: knowing np4 = -p4, the compiler turns
        // the select of s or s + np4 into a subtraction with a borrow chain
        // (VCC carry plus a hazard wait); p is uniform, so SGPRs are exact
        asm volatile("" : "+s"(np), "+s"(np2), "+s"(np4), "+s"(np8));
    }
};

// x >= c ? x - c : x, with nc = -c: the candidate is formed by one add and
// selected, so no borrow chain feeds the select
template <typename W>
__device__ __forceinline__ W csub(W x, W c, W nc)
{
    const W t = x + nc;
    return x >= c ? t : x;
}

// B p for B in {1, 2, 4, 8}
template <int B, typename W>
__device__ __forceinline__ W mult_p(const Bfly<W> &k)
{
    static_assert(B == 1 || B == 2 || B == 4 || B == 8, "multiple of p");
    return B == 1 ? k.p : B == 2 ? k.p2 : B == 4 ? k.p4 : k.p8;
}

// a value below B p brought below 8p (B <= 16): one conditional subtraction
// of 8p where B exceeds 8
template <int B, typename W>
__device__ __forceinline__ W red8(W s, const Bfly<W> &k)
{
    static_assert(B <= 16, "lazy bound");
    if constexpr (B > 8)
        return csub(s, k.p8, k.np8);
    return s;
}

// x*w mod p in [0, 4p) for any 64-bit x: q' = xh*qh + hi(xl*qh) + hi(xh*ql)
// is the Shoup quotient minus 0, 1 or 2, and x*w - q'*p = x*w + q'*(-p) mod 2^64.
__device__ __forceinline__ uint64_t shoup4(uint64_t x, uint64_t w, uint64_t wq, uint64_t np)
{
    const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32);
    const uint32_t ql = (uint32_t)wq, qh = (uint32_t)(wq >> 32);
    uint64_t q = (uint64_t)xh * qh + __umulhi(xl, qh);
    q += __umulhi(xh, ql);
    const uint32_t q0 = (uint32_t)q, q1 = (uint32_t)(q >> 32);
    const uint32_t w0 = (uint32_t)w, w1 = (uint32_t)(w >> 32);
    const uint32_t n0 = (uint32_t)np, n1 = (uint32_t)(np >> 32);
    uint64_t r = (uint64_t)xl * w0;
    r += (uint64_t)q0 * n0;
    const uint32_t hi = (uint32_t)(r >> 32) + xl * w1 + xh * w0 + q0 * n1 + q1 * n0;
    return ((uint64_t)hi << 32) | (uint32_t)r;
}

// Lazy forward rule.  Bound (in multiples of p) of register j of a forward
// pass of R stages, before the stage on local bit lb (lb = -1: after the
// pass), for inputs below b0 p.  A stage pairs j0 and j0 + 2^lb with the
// bound b of the larger input: the sum is below 2b p (brought below 8p when
// 2b > 8), the difference a - c + b p goes through the Shoup product (below
// 4p), except for twiddle 1 (in the last pass, j0 = 0 mod 2^lb), where it is
// kept like a sum.
constexpr int fwd_bound(int j, int lb, int R, int b0, bool last)
{
    if (lb == R - 1)
        return b0;
    const int pb = lb + 1;
    const int j0 = j & ~(1 << pb);
    const int a = fwd_bound(j0, pb, R, b0, last);
    const int c = fwd_bound(j0 | (1 << pb), pb, R, b0, last);
    const int b = a > c ? a : c;
    const bool prod = ((j >> pb) & 1) && !(last && (j0 & ((1 << pb) - 1)) == 0);
    return prod ? 4 : (2 * b > 8 ? 8 : 2 * b);
}

constexpr int fwd_pass_out(int R, int b0, bool last)
{
    int m = 0;
    for (int j = 0; j < (1 << R); ++j) {
        const int b = fwd_bound(j, -1, R, b0, last);
        m = b > m ? b : m;
    }
    return m;
}

// Lazy inverse rule.  Bound of register j of an inverse pass before the
// stage on local bit lb (stages run lb = 0 .. R-1; lb = R: after the pass).
// A stage pairs j0 (upper, bound a) and j0 + 2^lb (bound c): v = c w is
// below 4p, or for twiddle 1 (first inverse pass, j0 = 0 mod 2^lb) v = c
// brought below 8p; with V the matching multiple (4 or 8), the outputs
// a + v and a - v + V p are below (a' + V) p, a' = a unless a + V > 16, in
// which case a is first brought below 8p.
constexpr int inv_v(int c, bool triv)
{
    return triv ? (c <= 4 ? 4 : 8) : 4;
}

constexpr int inv_a(int a, int c, bool triv)
{
    return a + inv_v(c, triv) > 16 ? 8 : a;
}

constexpr int inv_bound(int j, int lb, int b0, bool first)
{
    if (lb == 0)
        return b0;
    const int pb = lb - 1;
    const int j0 = j & ~(1 << pb);
    const int a = inv_bound(j0, pb, b0, first);
    const int c = inv_bound(j0 | (1 << pb), pb, b0, first);
    const bool triv = first && (j0 & ((1 << pb) - 1)) == 0;
    return inv_a(a, c, triv) + inv_v(c, triv);
}

constexpr int inv_pass_out(int R, int b0, bool first)
{
    int m = 0;
    for (int j = 0; j < (1 << R); ++j) {
        const int b = inv_bound(j, R, b0, first);
        m = b > m ? b : m;
    }
    return m;
}

// Strict butterflies ([0, 2p) in and out).
template <typename W>
__device__ __forceinline__ void bfly_dif_strict(W &a, W &c, W w, W wq, const Bfly<W> &k)
{
    const W u = a, v = c;
    a = add_2p<W>(u, v, k.p2);
    c = shoup_lazy(sub_2p<W>(u, v, k.p2), w, wq, k.p);
}

template <typename W>
__device__ __forceinline__ void bfly_dif1_strict(W &a, W &c, const Bfly<W> &k)
{
    const W u = a, v = c;
    a = add_2p<W>(u, v, k.p2);
    const W d = u + k.p2 - v;
    c = d >= k.p2 ? d - k.p2 : d;
}

template <typename W>
__device__ __forceinline__ void bfly_dit_strict(W &a, W &c, W w, W wq, const Bfly<W> &k)
{
    const W u = a;
    const W v = shoup_lazy(c, w, wq, k.p);
    a = add_2p<W>(u, v, k.p2);
    const W d = u + k.p2 - v;
    c = d >= k.p2 ? d - k.p2 : d;
}

// Lazy forward butterfly on inputs below B p (B <= 8); TRIV: twiddle 1.
template <int B, bool TRIV>
__device__ __forceinline__ void bfly_dif_lazy(uint64_t &a, uint64_t &c, uint64_t w, uint64_t wq,
                                              const Bfly<uint64_t> &k)
{
    const uint64_t d = a - c + mult_p<B>(k);  // in (0, 2B p)
    a = red8<2 * B>(a + c, k);
    if constexpr (TRIV)
        c = red8<2 * B>(d, k);
    else
        c = shoup4(d, w, wq, k.np);
}

// Lazy inverse butterfly on a below BA p and c below BC p.
template <int BA, int BC, bool TRIV>
__device__ __forceinline__ void bfly_dit_lazy(uint64_t &a, uint64_t &c, uint64_t w, uint64_t wq,
                                              const Bfly<uint64_t> &k)
{
    constexpr int V = inv_v(BC, TRIV);
    const uint64_t v = TRIV ? red8<BC>(c, k) : shoup4(c, w, wq, k.np);
    const uint64_t ar = BA + V > 16 ? red8<BA>(a, k) : a;
    a = ar + v;
    c = ar - v + mult_p<V>(k);
}

// [0, 2p) (strict) or below B p (lazy, B <= 8) -> [0, p)
template <typename W, int B>
__device__ __forceinline__ W bfly_final(W x, const Bfly<W> &k)
{
    if constexpr (B > 4)
        x = csub(x, k.p4, k.np4);
    if constexpr (B > 2)
        x = csub(x, k.p2, k.np2);
    return csub(x, k.p, k.np);
}

template <int L, int LOGE, int S0>
struct PassShape {
    static constexpr int R = (L - S0 < LOGE) ? (L - S0) : LOGE;
    static constexpr int G = 1 << (LOGE - R);  // groups per thread
    static constexpr int T = 1 << (L - LOGE);  // threads
    static constexpr int BLO = L - S0 - R;
    __device__ static __forceinline__ int index(int t, int g, int j)
    {
        const int gi = t + T * g;
        const int glo = gi & ((1 << BLO) - 1);
        const int ghi = gi >> BLO;
        return (ghi << (BLO + R)) | (j << BLO) | glo;
    }
};

template <typename W, int L, int LOGE, int S0>
__device__ __forceinline__ void lds_store(const W (&x)[1 << LOGE], W *lds, int t)
{
    using PS = PassShape<L, LOGE, S0>;
#pragma unroll
    for (int g = 0; g < PS::G; ++g)
#pragma unroll
        for (int j = 0; j < (1 << PS::R); ++j)
            lds[swz(PS::index(t, g, j))] = x[g * (1 << PS::R) + j];
}

template <typename W, int L, int LOGE, int S0>
__device__ __forceinline__ void lds_load(W (&x)[1 << LOGE], const W *lds, int t)
{
    using PS = PassShape<L, LOGE, S0>;
#pragma unroll
    for (int g = 0; g < PS::G; ++g)
#pragma unroll
        for (int j = 0; j < (1 << PS::R); ++j)
            x[g * (1 << PS::R) + j] = lds[swz(PS::index(t, g, j))];
}

// Twiddles of one pass, loaded up front: stage with local bit LB uses the
// 2^LB values at flat positions 2^LB - 1 + jl (jl < 2^LB), as (value, Shoup)
// pairs.  They do not depend on the group (BLO <= log2 T), only on t.
template <typename W, int L, int LOGE, int S0>
__device__ __forceinline__ void load_pass_tw(W (&tw)[2 * ((1 << LOGE) - 1)], int t,
                                             const W *__restrict__ stab)
{
    using PS = PassShape<L, LOGE, S0>;
    constexpr int R = PS::R;
    const int glo = t & ((1 << PS::BLO) - 1);
#pragma unroll
    for (int f = 0; f < (1 << R) - 1; ++f) {
        const int LB = 31 - __builtin_clz(f + 1);
        const int jl = f + 1 - (1 << LB);
        const int s = S0 + (R - 1 - LB);
        const int off = (1 << L) - (1 << (L - s));
        load_tw<W>(stab, off + ((jl << PS::BLO) | glo), tw[2 * f], tw[2 * f + 1]);
    }
}

// forward (DIF) pair (J, J + 2^LB) of every group at the stage on local bit
// LB of the pass at S0, whose registers were below B0 p at the pass start
template <typename W, bool LOOSE, int L, int LOGE, int S0, int B0, int LB, int J>
__device__ __forceinline__ void fwd_pairs(W (&x)[1 << LOGE], const W (&tw)[2 * ((1 << LOGE) - 1)],
                                          const Bfly<W> &k)
{
    using PS = PassShape<L, LOGE, S0>;
    constexpr int R = PS::R;
    if constexpr ((J & (1 << LB)) == 0) {
        constexpr int jl = J & ((1 << LB) - 1);
        // in the last pass (BLO = 0) the stage-table index is jl itself, and
        // entry 0 of every stage table is w^0 = 1
        constexpr bool TRIV = PS::BLO == 0 && jl == 0;
        const W w = tw[2 * ((1 << LB) - 1 + jl)], wq = tw[2 * ((1 << LB) - 1 + jl) + 1];
#pragma unroll
        for (int g = 0; g < PS::G; ++g) {
            W &a = x[g * (1 << R) + J];
            W &c = x[g * (1 << R) + J + (1 << LB)];
            if constexpr (LOOSE) {
                constexpr int BA = fwd_bound(J, LB, R, B0, PS::BLO == 0);
                constexpr int BC = fwd_bound(J + (1 << LB), LB, R, B0, PS::BLO == 0);
                bfly_dif_lazy<(BA > BC ? BA : BC), TRIV>(a, c, w, wq, k);
            } else if constexpr (TRIV) {
                bfly_dif1_strict<W>(a, c, k);
            } else {
                bfly_dif_strict<W>(a, c, w, wq, k);
            }
        }
    }
    if constexpr (J + 1 < (1 << R))
        fwd_pairs<W, LOOSE, L, LOGE, S0, B0, LB, J + 1>(x, tw, k);
}

// forward stages of the pass at S0, local bit LB down to 0
template <typename W, bool LOOSE, int L, int LOGE, int S0, int B0, int LB>
__device__ __forceinline__ void fwd_stages(W (&x)[1 << LOGE], const W (&tw)[2 * ((1 << LOGE) - 1)],
                                           const Bfly<W> &k)
{
    fwd_pairs<W, LOOSE, L, LOGE, S0, B0, LB, 0>(x, tw, k);
    if constexpr (LB > 0)
        fwd_stages<W, LOOSE, L, LOGE, S0, B0, LB - 1>(x, tw, k);
}

// First stage of the first pass when the upper half of the input is zero:
// the pair (a, 0) becomes (a, a w).  The outputs stay within the bounds of
// the general stage (a below B0 p, a w below 4p, or 2p in strict mode).
template <typename W, bool LOOSE, int L, int LOGE, int J>
__device__ __forceinline__ void fwd_half_pairs(W (&x)[1 << LOGE],
                                               const W (&tw)[2 * ((1 << LOGE) - 1)],
                                               const Bfly<W> &k)
{
    using PS = PassShape<L, LOGE, 0>;
    constexpr int LB = PS::R - 1;
    constexpr bool TRIV = PS::BLO == 0 && J == 0;
    const W w = tw[2 * ((1 << LB) - 1 + J)], wq = tw[2 * ((1 << LB) - 1 + J) + 1];
#pragma unroll
    for (int g = 0; g < PS::G; ++g) {
        const W a = x[g * (1 << PS::R) + J];
        W &c = x[g * (1 << PS::R) + J + (1 << LB)];
        if constexpr (TRIV)
            c = a;
        else if constexpr (LOOSE)
            c = shoup4(a, w, wq, k.np);
        else
            c = shoup_lazy(a, w, wq, k.p);
    }
    if constexpr (J + 1 < (1 << LB))
        fwd_half_pairs<W, LOOSE, L, LOGE, J + 1>(x, tw, k);
}

// inverse (DIT) pair (J, J + 2^LB) of every group, stage LB of the pass at
// S0, whose registers were below B0 p at the pass start
template <typename W, bool LOOSE, int L, int LOGE, int S0, int B0, int LB, int J>
__device__ __forceinline__ void inv_pairs(W (&x)[1 << LOGE], const W (&tw)[2 * ((1 << LOGE) - 1)],
                                          const Bfly<W> &k)
{
    using PS = PassShape<L, LOGE, S0>;
    constexpr int R = PS::R;
    if constexpr ((J & (1 << LB)) == 0) {
        constexpr int jl = J & ((1 << LB) - 1);
        constexpr bool TRIV = PS::BLO == 0 && jl == 0;  // twiddle w^0 = 1 (see fwd_pairs)
        const W w = tw[2 * ((1 << LB) - 1 + jl)], wq = tw[2 * ((1 << LB) - 1 + jl) + 1];
#pragma unroll
        for (int g = 0; g < PS::G; ++g) {
            W &a = x[g * (1 << R) + J];
            W &c = x[g * (1 << R) + J + (1 << LB)];
            if constexpr (LOOSE) {
                constexpr int BA = inv_bound(J, LB, B0, PS::BLO == 0);
                constexpr int BC = inv_bound(J + (1 << LB), LB, B0, PS::BLO == 0);
                bfly_dit_lazy<BA, BC, TRIV>(a, c, w, wq, k);
            } else if constexpr (TRIV) {
                bfly_dif1_strict<W>(a, c, k);  // (a + c, a - c) either way
            } else {
                bfly_dit_strict<W>(a, c, w, wq, k);
            }
        }
    }
    if constexpr (J + 1 < (1 << R))
        inv_pairs<W, LOOSE, L, LOGE, S0, B0, LB, J + 1>(x, tw, k);
}

// inverse stages of the pass at S0, local bit LB up to R-1
template <typename W, bool LOOSE, int L, int LOGE, int S0, int B0, int LB>
__device__ __forceinline__ void inv_stages(W (&x)[1 << LOGE], const W (&tw)[2 * ((1 << LOGE) - 1)],
                                           const Bfly<W> &k)
{
    inv_pairs<W, LOOSE, L, LOGE, S0, B0, LB, 0>(x, tw, k);
    if constexpr (LB + 1 < PassShape<L, LOGE, S0>::R)
        inv_stages<W, LOOSE, L, LOGE, S0, B0, LB + 1>(x, tw, k);
}

// forward passes from S0 to the end (x holds the pass-S0 arrangement and tw
// its twiddles, registers below B0 p); `half`: the upper half of the input
// is zero (first pass only).  On return the last pass's registers hold runs
// of consecutive positions, below fwd_bound(j, -1, R, fwd_last_b0) p.
template <typename W, bool LOOSE, int L, int LOGE, int S0, int B0>
__device__ __forceinline__ void fwd_passes(W (&x)[1 << LOGE], W (&tw)[2 * ((1 << LOGE) - 1)],
                                           W *lds, int t, const W *__restrict__ stw,
                                           const Bfly<W> &k, bool half, bool trunc)
{
    using PS = PassShape<L, LOGE, S0>;
    if constexpr (S0 == 0) {
        if (half)
            fwd_half_pairs<W, LOOSE, L, LOGE, 0>(x, tw, k);
        else
            fwd_pairs<W, LOOSE, L, LOGE, 0, B0, PS::R - 1, 0>(x, tw, k);
        if constexpr (PS::R > 1)
            fwd_stages<W, LOOSE, L, LOGE, 0, B0, PS::R - 2>(x, tw, k);
    } else {
        // later passes act below the index's top two bits, which come from
        // the thread (one group per thread): a thread whose positions all
        // lie in the dropped quarter (whole waves for most shapes) skips.
        // Not with 16 words per thread: the branch takes the kernel past
        // the 128 registers of four waves per SIMD.
        bool skip = false;
        if constexpr (PS::G == 1 && LOGE >= 5)
            skip = trunc && PS::index(t, 0, 0) >= (3 << (L - 2));
        if (!skip)
            fwd_stages<W, LOOSE, L, LOGE, S0, B0, PS::R - 1>(x, tw, k);
    }
    if constexpr (S0 + PS::R < L) {
        load_pass_tw<W, L, LOGE, S0 + PS::R>(tw, t, stw);  // in flight during the exchange
        lds_store<W, L, LOGE, S0>(x, lds, t);
        __syncthreads();
        lds_load<W, L, LOGE, S0 + PS::R>(x, lds, t);
        __syncthreads();
        constexpr int B1 = LOOSE ? fwd_pass_out(PS::R, B0, false) : 2;
        fwd_passes<W, LOOSE, L, LOGE, S0 + PS::R, B1>(x, tw, lds, t, stw, k, false, trunc);
    }
}

// first stage of the last forward pass (it ends at stage L)
template <int L, int LOGE>
constexpr int last_pass_start()
{
    return ((L - 1) / LOGE) * LOGE;
}

// register bound at the start of the last forward pass, for inputs below b0 p
template <int L, int LOGE>
constexpr int fwd_last_b0(int b0)
{
    int b = b0;
    for (int s0 = 0; s0 < last_pass_start<L, LOGE>(); s0 += LOGE)
        b = fwd_pass_out(LOGE, b, false);
    return b;
}

// lazy registers of the last pass (run length 2^R) brought to [0, p)
template <typename W, bool LOOSE, int LOGE, int R, int B0, int J>
__device__ __forceinline__ void fwd_final(const W (&x)[1 << LOGE], int base, W *y, const Bfly<W> &k)
{
    constexpr int B = LOOSE ? fwd_bound(J, -1, R, B0, true) : 2;
    y[J] = bfly_final<W, B>(x[base + J], k);
    if constexpr (J + 1 < (1 << R))
        fwd_final<W, LOOSE, LOGE, R, B0, J + 1>(x, base, y, k);
}

// CNT consecutive words (CNT a power of two, 16-byte aligned runs when
// CNT * sizeof(W) >= 16) moved with vector loads / stores
template <typename W, int CNT>
__device__ __forceinline__ void store_words(W *dst, const W *v)
{
    if constexpr (CNT * sizeof(W) >= 16) {
        constexpr int PER = 16 / sizeof(W);
#pragma unroll
        for (int q = 0; q < CNT / PER; ++q) {
            uint4 u;
            if constexpr (sizeof(W) == 8) {
                u.x = (uint32_t)v[2 * q];
                u.y = (uint32_t)(v[2 * q] >> 32);
                u.z = (uint32_t)v[2 * q + 1];
                u.w = (uint32_t)(v[2 * q + 1] >> 32);
            } else {
                u.x = v[4 * q];
                u.y = v[4 * q + 1];
                u.z = v[4 * q + 2];
                u.w = v[4 * q + 3];
            }
            reinterpret_cast<uint4 *>(dst)[q] = u;
        }
    } else {
#pragma unroll
        for (int j = 0; j < CNT; ++j)
            dst[j] = v[j];
    }
}

template <typename W, int CNT>
__device__ __forceinline__ void load_words(const W *src, W *v)
{
    if constexpr (CNT * sizeof(W) >= 16) {
        constexpr int PER = 16 / sizeof(W);
#pragma unroll
        for (int q = 0; q < CNT / PER; ++q) {
            const uint4 u = reinterpret_cast<const uint4 *>(src)[q];
            if constexpr (sizeof(W) == 8) {
                v[2 * q] = ((uint64_t)u.y << 32) | u.x;
                v[2 * q + 1] = ((uint64_t)u.w << 32) | u.z;
            } else {
                v[4 * q] = u.x;
                v[4 * q + 1] = u.y;
                v[4 * q + 2] = u.z;
                v[4 * q + 3] = u.w;
            }
        }
    } else {
#pragma unroll
        for (int j = 0; j < CNT; ++j)
            v[j] = src[j];
    }
}

// inverse passes from S0 down to 0 (registers below B0 p at the start of
// the pass at S0); on return x holds the pass-0 arrangement, below
// inv_pass_out(...) p in lazy mode (any 64-bit value is a valid input of the
// final scaling or of the merge pass's twist)
// trunc (3/4-truncated product): passes above the last skip the threads
// whose positions all lie in the dropped quarter (their values only reach
// that quarter's outputs, which the caller's correction uses whatever they
// are).
template <typename W, bool LOOSE, int L, int LOGE, int S0, int B0>
__device__ __forceinline__ void inv_passes(W (&x)[1 << LOGE], W (&tw)[2 * ((1 << LOGE) - 1)],
                                           W *lds, int t, const NttTab<W> &tab,
                                           const Bfly<W> &k, bool trunc)
{
    using PS = PassShape<L, LOGE, S0>;
    bool skip = false;
    if constexpr (S0 > 0 && PS::G == 1 && LOGE >= 5)  // (see fwd_passes)
        skip = trunc && PS::index(t, 0, 0) >= (3 << (L - 2));
    if (!skip)
        inv_stages<W, LOOSE, L, LOGE, S0, B0, 0>(x, tw, k);
    if constexpr (S0 > 0) {
        load_pass_tw<W, L, LOGE, S0 - LOGE>(tw, t, tab.sitw);
        __syncthreads();  // every thread has finished reading the previous arrangement
        lds_store<W, L, LOGE, S0>(x, lds, t);
        __syncthreads();
        lds_load<W, L, LOGE, S0 - LOGE>(x, lds, t);
        constexpr int B1 = LOOSE ? inv_pass_out(PS::R, B0, PS::BLO == 0) : 2;
        inv_passes<W, LOOSE, L, LOGE, S0 - LOGE, B1>(x, tw, lds, t, tab, k, trunc);
    }
}

// pre: the input rows are the column pass's outputs (row g of entry e at
// in + e*in_stride + g*M, every word valid): no radix-R step on the load.
template <typename Win, typename W, bool LOOSE, int L, int LOGE>
__device__ __forceinline__ void fwd_reg_body(const Win *__restrict__ in, int64_t in_stride,
                                             int64_t in_len, W *__restrict__ out, int64_t units,
                                             int64_t u, int logR, int logP, const NttTab<W> &tab,
                                             W p, bool pre = false, bool trunc = false)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    W *lds = reinterpret_cast<W *>(smem_raw);
    constexpr int E = 1 << LOGE, M = 1 << L;
    const int t = threadIdx.x;
    const int64_t e = u >> logR;
    const int R = 1 << logR;
    const int gg = (int)(u & (R - 1));
    const int logN = L + logR;
    const int64_t N = (int64_t)1 << logN;
    const W p2 = p + p;
    const Win *xin = in + e * in_stride;
    // truncated: with a radix-4 split the fourth sub-transform (frequencies
    // 3 mod 4) is dropped (the whole workgroup returns), without a split the
    // last quarter of the positions
    if (trunc && R == 4 && gg == 3)
        return;
    const bool tq = trunc && R == 1;

    W tw[2 * (E - 1)];
    load_pass_tw<W, L, LOGE, 0>(tw, t, tab.stw);
    W x[E];
    // pass-0 arrangement: element j of thread t is index (j << (L-LOGE)) | t
#pragma unroll
    for (int j = 0; j < E; ++j) {
        const int i1 = (j << (L - LOGE)) | t;
        W v = 0;
        if (pre) {
            v = (W)xin[(int64_t)gg * M + i1];
        } else if (R == 1) {
            if (i1 < in_len)
                v = (W)xin[i1];
        } else {
            W acc = 0;
            for (int t2 = 0; t2 < R; ++t2) {
                const int64_t src = (int64_t)i1 + (int64_t)M * t2;
                if (src < in_len) {
                    const int64_t widx = ((int64_t)M * gg * t2) & (N - 1);
                    acc = add_2p<W>(acc, shoup_lazy((W)xin[src], tab.tw[widx], tab.twq[widx], p), p2);
                }
            }
            const int64_t widx = (int64_t)i1 * gg;
            v = shoup_lazy(acc, tab.tw[widx], tab.twq[widx], p);
        }
        x[j] = v;
    }
    const Bfly<W> k(p);
    // inputs below 2p (coefficients, the column-in-load sums, or the column
    // pass's reduced rows); a direct load of at most M/2 coefficients leaves
    // the upper half zero
    fwd_passes<W, LOOSE, L, LOGE, 0, 2>(x, tw, lds, t, tab.stw, k,
                                        !pre && R == 1 && in_len <= (M >> 1), tq);
    // the last pass leaves each thread runs of 2^R consecutive positions:
    // store them from registers (no LDS round trip)
    const int64_t count = units >> logR;
    using PSL = PassShape<L, LOGE, last_pass_start<L, LOGE>()>;
    constexpr int RUN = 1 << PSL::R;
    // words per vector store: within one layout block of 2^logP positions
    // (logP >= 2 unless N < 4)
    constexpr int VEC = RUN < 4 ? RUN : 4;
#pragma unroll
    for (int g = 0; g < PSL::G; ++g) {
        // a truncated transform keeps the positions below 3M/4 (a run lies
        // within one quarter)
        if (tq && PSL::index(t, g, 0) >= (3 << (L - 2)))
            continue;
        W y[RUN];
        fwd_final<W, LOOSE, LOGE, PSL::R, fwd_last_b0<L, LOGE>(2), 0>(x, g * RUN, y, k);
        if (logP >= 2) {
#pragma unroll
            for (int j0 = 0; j0 < RUN; j0 += VEC)
                store_words<W, VEC>(out + blk_addr(e, (int64_t)gg * M + PSL::index(t, g, j0), count,
                                                   logP),
                                    y + j0);
        } else {
#pragma unroll
            for (int j = 0; j < RUN; ++j)
                out[blk_addr(e, (int64_t)gg * M + PSL::index(t, g, j), count, logP)] = y[j];
        }
    }
}

// TRUNC (3/4-truncated transforms) is a template parameter: its branches,
// even when not taken, take the radix-32 kernels past 256 registers (one
// wave per SIMD instead of two).
template <typename Win, typename W, bool LOOSE, int L, int LOGE, bool TRUNC>
__global__ void __launch_bounds__(1 << (L - LOGE)) __attribute__((amdgpu_waves_per_eu(LOGE == 5 ? kFwdWaves32 : 1)))
ntt_fwd_reg(const Win *__restrict__ in, int64_t in_stride, int64_t in_len, W *__restrict__ out,
            int64_t units, int logR, int logP, NttTab<W> tab, W p, int pre)
{
    // pre: the column pass's rows as input
    fwd_reg_body<Win, W, LOOSE, L, LOGE>(in, in_stride, in_len, out, units,
                                         xcd_unit(blockIdx.x, units), logR, logP, tab, p,
                                         pre != 0, TRUNC);
}

template <typename W, typename Wout, bool LOOSE, int L, int LOGE>
__device__ __forceinline__ void inv_reg_body(const W *__restrict__ in, Wout *__restrict__ out,
                                             int64_t out_stride, int64_t lo, int64_t hi,
                                             W *__restrict__ tmp, int64_t units, int64_t u,
                                             int logR, int logP, const NttTab<W> &tab, W p,
                                             bool trunc = false)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    W *lds = reinterpret_cast<W *>(smem_raw);
    constexpr int E = 1 << LOGE, M = 1 << L;
    const int t = threadIdx.x;
    const int64_t e = u >> logR;
    const int R = 1 << logR;
    const int gg = (int)(u & (R - 1));
    const int64_t N = (int64_t)M << logR;
    const int64_t count = units >> logR;
    constexpr int SLAST = ((L - 1) / LOGE) * LOGE;
    if (trunc && R == 4 && gg == 3)  // dropped sub-transform (see fwd_reg_body)
        return;
    const bool tq = trunc && R == 1;
    W tw[2 * (E - 1)];
    load_pass_tw<W, L, LOGE, SLAST>(tw, t, tab.sitw);
    // the first inverse pass (stages SLAST..L-1) works on runs of 2^R
    // consecutive positions per thread: load them straight from global memory
    using PSF = PassShape<L, LOGE, SLAST>;
    constexpr int RUN = 1 << PSF::R;
    constexpr int VEC = RUN < 4 ? RUN : 4;
    W x[E];
#pragma unroll
    for (int g = 0; g < PSF::G; ++g) {
        if (tq && PSF::index(t, g, 0) >= (3 << (L - 2))) {
            // dropped quarter of a truncated product: never read
#pragma unroll
            for (int j = 0; j < RUN; ++j)
                x[g * RUN + j] = 0;
            continue;
        }
        if (logP >= 2) {
#pragma unroll
            for (int j0 = 0; j0 < RUN; j0 += VEC)
                load_words<W, VEC>(in + blk_addr(e, (int64_t)gg * M + PSF::index(t, g, j0), count,
                                                 logP),
                                   x + g * RUN + j0);
        } else {
#pragma unroll
            for (int j = 0; j < RUN; ++j)
                x[g * RUN + j] = in[blk_addr(e, (int64_t)gg * M + PSF::index(t, g, j), count, logP)];
        }
    }
    const Bfly<W> k(p);
    // inputs below 8p (the pointwise products' lazy outputs)
    inv_passes<W, LOOSE, L, LOGE, SLAST, 8>(x, tw, lds, t, tab, k, tq);
    // pass-0 arrangement: element j is index (j << (L-LOGE)) | t
    if (R == 1 && tq) {
        // truncated: scaled outputs corrected by the dropped quarter
        // (trunc_correct); register j + q E/4 holds position i + q N/4
        Wout *o = out + e * out_stride;
        constexpr int E4 = E >> 2;
#pragma unroll
        for (int j = 0; j < E4; ++j) {
            W c0 = shoup<W>(x[j], tab.ninv, tab.ninvq, p);
            W c1 = shoup<W>(x[j + E4], tab.ninv, tab.ninvq, p);
            W c2 = shoup<W>(x[j + 2 * E4], tab.ninv, tab.ninvq, p);
            trunc_correct<W>(c0, c1, c2, shoup<W>(x[j + 3 * E4], tab.ninv, tab.ninvq, p), tab, p);
            const int i1 = (j << (L - LOGE)) | t;
            const int q1 = M >> 2;
            if (i1 >= lo && i1 < hi)
                o[i1 - lo] = (Wout)c0;
            if (i1 + q1 >= lo && i1 + q1 < hi)
                o[i1 + q1 - lo] = (Wout)c1;
            if (i1 + 2 * q1 >= lo && i1 + 2 * q1 < hi)
                o[i1 + 2 * q1 - lo] = (Wout)c2;
        }
    } else if (R == 1) {
        Wout *o = out + e * out_stride;
#pragma unroll
        for (int j = 0; j < E; ++j) {
            const int i1 = (j << (L - LOGE)) | t;
            if (i1 >= lo && i1 < hi)
                o[i1 - lo] = (Wout)shoup<W>(x[j], tab.ninv, tab.ninvq, p);
        }
    } else {
        // lazy values: the merge pass starts with a Shoup product (twist),
        // whose output is in [0, 2p) for any word
        W *o = tmp + e * N + (int64_t)gg * M;
#pragma unroll
        for (int j = 0; j < E; ++j)
            o[(j << (L - LOGE)) | t] = x[j];
    }
}

// With 16 elements per thread the inverse is held to 128 VGPRs, so that four
// waves share each SIMD and hide the LDS-exchange latency.
template <typename W, typename Wout, bool LOOSE, int L, int LOGE>
__global__ void __launch_bounds__(1 << (L - LOGE)) __attribute__((amdgpu_waves_per_eu(LOGE == 4 ? kInvWaves : 1)))
ntt_inv_reg(const W *__restrict__ in, Wout *__restrict__ out, int64_t out_stride, int64_t lo,
            int64_t hi, W *__restrict__ tmp, int64_t units, int logR, int logP, NttTab<W> tab, W p,
            int trunc)
{
    inv_reg_body<W, Wout, LOOSE, L, LOGE>(in, out, out_stride, lo, hi, tmp, units,
                                          xcd_unit(blockIdx.x, units), logR, logP, tab, p,
                                          trunc != 0);
}

// Register-kernel launchers.  logE (elements per thread) is a tuning
// parameter: 4 (radix-16 passes, more threads, fewer registers) or 5
// (radix-32 passes, fewer LDS exchanges).
template <typename Win, typename W, bool LOOSE, int L, int LOGE, bool TRUNC>
hipError_t launch_fwd_reg4(const Win *in, int64_t in_stride, int64_t in_len, W *out,
                           int64_t units, int logR, int logP, const NttTab<W> &tab, W p,
                           hipStream_t s, int pre)
{
    constexpr int T = 1 << (L - LOGE);
    const size_t lds = (size_t)((1 << L) + (1 << L) / 32) * sizeof(W);
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)ntt_fwd_reg<Win, W, LOOSE, L, LOGE, TRUNC>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess)
            return e;
        attr = true;
    }
    hipLaunchKernelGGL((ntt_fwd_reg<Win, W, LOOSE, L, LOGE, TRUNC>), dim3((unsigned)units), dim3(T),
                       lds, s, in, in_stride, in_len, out, units, logR, logP, tab, p, pre);
    return hipGetLastError();
}

// pre: bit 0 the column pass's rows as input, bit 1 a 3/4-truncated transform
template <typename Win, typename W, bool LOOSE, int L, int LOGE>
hipError_t launch_fwd_reg3(const Win *in, int64_t in_stride, int64_t in_len, W *out,
                           int64_t units, int logR, int logP, const NttTab<W> &tab, W p,
                           hipStream_t s, int pre)
{
    if (pre & 2)
        return launch_fwd_reg4<Win, W, LOOSE, L, LOGE, true>(in, in_stride, in_len, out, units,
                                                             logR, logP, tab, p, s, pre & 1);
    return launch_fwd_reg4<Win, W, LOOSE, L, LOGE, false>(in, in_stride, in_len, out, units, logR,
                                                          logP, tab, p, s, pre & 1);
}

template <typename W, typename Wout, bool LOOSE, int L, int LOGE>
hipError_t launch_inv_reg3(const W *in, Wout *out, int64_t out_stride, int64_t lo, int64_t hi,
                           W *tmp, int64_t units, int logR, int logP, const NttTab<W> &tab, W p,
                           hipStream_t s, int trunc)
{
    constexpr int T = 1 << (L - LOGE);
    const size_t lds = (size_t)((1 << L) + (1 << L) / 32) * sizeof(W);
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)ntt_inv_reg<W, Wout, LOOSE, L, LOGE>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess)
            return e;
        attr = true;
    }
    hipLaunchKernelGGL((ntt_inv_reg<W, Wout, LOOSE, L, LOGE>), dim3((unsigned)units), dim3(T), lds,
                       s, in, out, out_stride, lo, hi, tmp, units, logR, logP, tab, p, trunc);
    return hipGetLastError();
}

template <typename Win, typename W, bool LOOSE, int L>
hipError_t launch_fwd_reg2(int logE, const Win *in, int64_t in_stride, int64_t in_len, W *out,
                           int64_t units, int logR, int logP, const NttTab<W> &tab, W p,
                           hipStream_t s, int pre)
{
    if (logE == 5 && L >= 12)
        return launch_fwd_reg3<Win, W, LOOSE, L, (L >= 12 ? 5 : 4)>(in, in_stride, in_len, out,
                                                                   units, logR, logP, tab, p, s,
                                                                   pre);
    return launch_fwd_reg3<Win, W, LOOSE, L, 4>(in, in_stride, in_len, out, units, logR, logP, tab,
                                                p, s, pre);
}

template <typename W, typename Wout, bool LOOSE, int L>
hipError_t launch_inv_reg2(int logE, const W *in, Wout *out, int64_t out_stride, int64_t lo,
                           int64_t hi, W *tmp, int64_t units, int logR, int logP,
                           const NttTab<W> &tab, W p, hipStream_t s, int trunc)
{
    if (logE == 5 && L >= 12)
        return launch_inv_reg3<W, Wout, LOOSE, L, (L >= 12 ? 5 : 4)>(in, out, out_stride, lo, hi,
                                                                    tmp, units, logR, logP, tab,
                                                                    p, s, trunc);
    return launch_inv_reg3<W, Wout, LOOSE, L, 4>(in, out, out_stride, lo, hi, tmp, units, logR,
                                                 logP, tab, p, s, trunc);
}

template <typename Win, typename W, int L>
hipError_t launch_fwd_reg(int logE, const Win *in, int64_t in_stride, int64_t in_len, W *out,
                          int64_t units, int logR, int logP, const NttTab<W> &tab, W p,
                          hipStream_t s, int pre)
{
    if constexpr (sizeof(W) == 8) {
        if ((uint64_t)p < ((uint64_t)1 << 60))
            return launch_fwd_reg2<Win, W, true, L>(logE, in, in_stride, in_len, out, units, logR,
                                                    logP, tab, p, s, pre);
    }
    return launch_fwd_reg2<Win, W, false, L>(logE, in, in_stride, in_len, out, units, logR, logP,
                                             tab, p, s, pre);
}

template <typename W, typename Wout, int L>
hipError_t launch_inv_reg(int logE, const W *in, Wout *out, int64_t out_stride, int64_t lo,
                          int64_t hi, W *tmp, int64_t units, int logR, int logP,
                          const NttTab<W> &tab, W p, hipStream_t s, int trunc)
{
    if constexpr (sizeof(W) == 8) {
        if ((uint64_t)p < ((uint64_t)1 << 60))
            return launch_inv_reg2<W, Wout, true, L>(logE, in, out, out_stride, lo, hi, tmp, units,
                                                     logR, logP, tab, p, s, trunc);
    }
    return launch_inv_reg2<W, Wout, false, L>(logE, in, out, out_stride, lo, hi, tmp, units, logR,
                                              logP, tab, p, s, trunc);
}

template <typename Win, typename W>
hipError_t dispatch_fwd_reg(int logM, int logE, const Win *in, int64_t in_stride, int64_t in_len,
                            W *out, int64_t units, int logR, int logP, const NttTab<W> &tab, W p,
                            hipStream_t s, int pre = 0)
{
    switch (logM) {
    case 11:
        return launch_fwd_reg<Win, W, 11>(logE, in, in_stride, in_len, out, units, logR, logP, tab,
                                          p, s, pre);
    case 12:
        return launch_fwd_reg<Win, W, 12>(logE, in, in_stride, in_len, out, units, logR, logP, tab,
                                          p, s, pre);
    case 13:
        return launch_fwd_reg<Win, W, 13>(logE, in, in_stride, in_len, out, units, logR, logP, tab,
                                          p, s, pre);
    case 14:
        if constexpr (sizeof(W) == 4)
            return launch_fwd_reg<Win, W, 14>(logE, in, in_stride, in_len, out, units, logR, logP,
                                              tab, p, s, pre);
        [[fallthrough]];
    default:
        return hipErrorInvalidValue;
    }
}

template <typename W, typename Wout>
hipError_t dispatch_inv_reg(int logM, int logE, const W *in, Wout *out, int64_t out_stride,
                            int64_t lo, int64_t hi, W *tmp, int64_t units, int logR, int logP,
                            const NttTab<W> &tab, W p, hipStream_t s, int trunc = 0)
{
    switch (logM) {
    case 11:
        return launch_inv_reg<W, Wout, 11>(logE, in, out, out_stride, lo, hi, tmp, units, logR,
                                           logP, tab, p, s, trunc);
    case 12:
        return launch_inv_reg<W, Wout, 12>(logE, in, out, out_stride, lo, hi, tmp, units, logR,
                                           logP, tab, p, s, trunc);
    case 13:
        return launch_inv_reg<W, Wout, 13>(logE, in, out, out_stride, lo, hi, tmp, units, logR,
                                           logP, tab, p, s, trunc);
    case 14:
        if constexpr (sizeof(W) == 4)
            return launch_inv_reg<W, Wout, 14>(logE, in, out, out_stride, lo, hi, tmp, units, logR,
                                               logP, tab, p, s, trunc);
        [[fallthrough]];
    default:
        return hipErrorInvalidValue;
    }
}

// ===========================================================================
// Column pass of the long transforms, N = R M with R > 4 (M the LDS limit).
// X[R k + g] = sum_i1 w_M^(i1 k) y_g[i1] with y_g[i1] = w^(i1 g) sum_t
// x[i1 + M t] w_R^(t g): the column pass forms every y_g (one R-point
// transform per column i1, then the twist w^(i1 g)) into rows of M words,
// and the register kernels transform the rows.  The inverse runs the rows'
// M-point inverses into tmp, then x[i1 + M t] = N^-1 sum_g w_R^(-t g)
// w^(-i1 g) z_g[i1] here.  One workgroup holds 2^logC consecutive columns
// of one entry (all R rows) in LDS, [t][c], values in [0, 2p).
// ===========================================================================
constexpr int kColThreads = 256;

__device__ __forceinline__ int bitrev(int j, int bits)
{
    return (int)(__builtin_bitreverse32((uint32_t)j) >> (32 - bits));
}

// w^a for a < N from the two-level tables, times x in [0, 2p): [0, 2p)
template <typename W>
__device__ __forceinline__ W twist(W x, int64_t a, const W *lo, const W *loq, const W *hi,
                                   const W *hiq, int slog, W p)
{
    const int64_t al = a & (((int64_t)1 << slog) - 1), ah = a >> slog;
    return shoup_lazy(shoup_lazy(x, lo[al], loq[al], p), hi[ah], hiq[ah], p);
}

template <typename Win, typename W>
__global__ void __launch_bounds__(kColThreads)
ntt_cols_fwd(const Win *__restrict__ in, int64_t in_stride, int64_t in_len, Win *__restrict__ tmp,
             int logM, int logR, int logC, ColTab<W> ct, W p)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    W *s = reinterpret_cast<W *>(smem_raw);
    const int C = 1 << logC, R = 1 << logR;
    const int64_t M = (int64_t)1 << logM, N = M << logR;
    const int64_t groups = M >> logC;  // column groups per entry
    const int64_t e = blockIdx.x / groups;
    const int64_t c0 = (blockIdx.x - e * groups) << logC;
    const Win *x = in + e * in_stride;
    const W p2 = p + p;
    for (int idx = threadIdx.x; idx < R * C; idx += kColThreads) {
        const int64_t src = c0 + (idx & (C - 1)) + M * (idx >> logC);
        s[idx] = src < in_len ? (W)x[src] : (W)0;
    }
    __syncthreads();
    // R-point decimation in frequency over the rows: natural in, slot j
    // holds output g = bitrev(j) afterwards
    for (int lh = logR - 1; lh >= 0; --lh) {
        const int h = 1 << lh;
        for (int b = threadIdx.x; b < (R / 2) * C; b += kColThreads) {
            const int c = b & (C - 1), j = b >> logC;
            const int jj = j & (h - 1);
            const int i0 = (((j >> lh) << (lh + 1)) + jj) * C + c;
            const int i1 = i0 + h * C;
            const int k = jj << (logR - 1 - lh);
            const W u = s[i0], v = s[i1];
            s[i0] = add_2p<W>(u, v, p2);
            s[i1] = shoup_lazy(sub_2p<W>(u, v, p2), ct.rt[k], ct.rtq[k], p);
        }
        __syncthreads();
    }
    Win *y = tmp + e * N;
    for (int idx = threadIdx.x; idx < R * C; idx += kColThreads) {
        const int c = idx & (C - 1);
        const int g = bitrev(idx >> logC, logR);
        const int64_t i1 = c0 + c;
        const W v = twist<W>(s[idx], i1 * g, ct.lo, ct.loq, ct.hi, ct.hiq, ct.slog, p);
        y[(int64_t)g * M + i1] = (Win)reduce_2p<W>(v, p);
    }
}

template <typename W, typename Wout>
__global__ void __launch_bounds__(kColThreads)
ntt_cols_inv(const W *__restrict__ tmp, Wout *__restrict__ out, int64_t out_stride, int64_t lo,
             int64_t hi, int logM, int logR, int logC, ColTab<W> ct, W ninv, W ninvq, W p)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    W *s = reinterpret_cast<W *>(smem_raw);
    const int C = 1 << logC, R = 1 << logR;
    const int64_t M = (int64_t)1 << logM, N = M << logR;
    const int64_t groups = M >> logC;
    const int64_t e = blockIdx.x / groups;
    const int64_t c0 = (blockIdx.x - e * groups) << logC;
    const W *z = tmp + e * N;
    const W p2 = p + p;
    // slot j <- w^(-i1 g) z_g[i1] with g = bitrev(j): bit-reversed order in
    for (int idx = threadIdx.x; idx < R * C; idx += kColThreads) {
        const int c = idx & (C - 1);
        const int g = bitrev(idx >> logC, logR);
        const int64_t i1 = c0 + c;
        s[idx] = twist<W>(z[(int64_t)g * M + i1], i1 * g, ct.loi, ct.loiq, ct.hii, ct.hiiq, ct.slog,
                          p);
    }
    __syncthreads();
    // R-point decimation in time with w_R^-1: natural order out
    for (int lh = 0; lh < logR; ++lh) {
        const int h = 1 << lh;
        for (int b = threadIdx.x; b < (R / 2) * C; b += kColThreads) {
            const int c = b & (C - 1), j = b >> logC;
            const int jj = j & (h - 1);
            const int i0 = (((j >> lh) << (lh + 1)) + jj) * C + c;
            const int i1 = i0 + h * C;
            const int k = jj << (logR - 1 - lh);
            const W u = s[i0];
            const W v = shoup_lazy(s[i1], ct.rti[k], ct.rtiq[k], p);
            s[i0] = add_2p<W>(u, v, p2);
            const W d = sub_2p<W>(u, v, p2);
            s[i1] = d >= p2 ? d - p2 : d;
        }
        __syncthreads();
    }
    Wout *o = out + e * out_stride;
    for (int idx = threadIdx.x; idx < R * C; idx += kColThreads) {
        const int64_t pos = c0 + (idx & (C - 1)) + M * (idx >> logC);
        if (pos >= lo && pos < hi)
            o[pos - lo] = (Wout)shoup<W>(s[idx], ninv, ninvq, p);
    }
}

// columns per workgroup: 16 (128-byte runs of 64-bit words) while R x 16
// words fit 64 KiB of LDS (the default limit of dynamic LDS), fewer for the
// largest R
template <typename W>
int col_log_c(int logR)
{
    const int room = 16 - (sizeof(W) == 8 ? 3 : 2) - logR;  // log2(65536 / (R * sizeof(W)))
    return room < 2 ? 2 : (room > 4 ? 4 : room);
}

template <typename Win, typename W>
hipError_t launch_cols_fwd(const Win *in, int64_t in_stride, int64_t in_len, Win *tmp,
                           int64_t count, int logM, int logR, const ColTab<W> &ct, W p,
                           hipStream_t s)
{
    const int logC = col_log_c<W>(logR);
    const size_t lds = ((size_t)1 << (logR + logC)) * sizeof(W);
    const int64_t blocks = count << (logM - logC);
    hipLaunchKernelGGL((ntt_cols_fwd<Win, W>), dim3((unsigned)blocks), dim3(kColThreads), lds, s, in,
                       in_stride, in_len, tmp, logM, logR, logC, ct, p);
    return hipGetLastError();
}

template <typename W, typename Wout>
hipError_t launch_cols_inv(const W *tmp, Wout *out, int64_t out_stride, int64_t lo, int64_t hi,
                           int64_t count, int logM, int logR, const ColTab<W> &ct, W ninv,
                           W ninvq, W p, hipStream_t s)
{
    const int logC = col_log_c<W>(logR);
    const size_t lds = ((size_t)1 << (logR + logC)) * sizeof(W);
    const int64_t blocks = count << (logM - logC);
    hipLaunchKernelGGL((ntt_cols_inv<W, Wout>), dim3((unsigned)blocks), dim3(kColThreads), lds, s,
                       tmp, out, out_stride, lo, hi, logM, logR, logC, ct, ninv, ninvq, p);
    return hipGetLastError();
}

constexpr int kRegMinLog = 11;

// Stage tables for sub-transforms of 2^logM points.  The tables of length
// 2^logN are built for min(logN, max_log_m); a truncated transform split in
// four one level below that (logN = max_log_m + 1) uses the larger table's
// stages from the second on: stage s of 2^logM points (root w^(N/M)) is
// stage s + 1 of 2^(logM + 1), which sits 2^logM pairs further on.
template <typename W>
NttTab<W> tab_for_sub(const NttTab<W> &tab, int logN, int logM)
{
    NttTab<W> t = tab;
    const int built = logN < NttLimits<W>::max_log_m ? logN : NttLimits<W>::max_log_m;
    if (logM < built) {
        t.stw += (size_t)2 << logM;
        t.sitw += (size_t)2 << logM;
    }
    return t;
}

}  // namespace

template <typename Win, typename W>
hipError_t launch_ntt_fwd(const Win *in, int64_t in_stride, int64_t in_len, W *out,
                          int64_t count, int logN, int logP, const NttTune &tune,
                          const NttTab<W> &tab, W p, hipStream_t s, Win *tmp,
                          const ColTab<W> *ct, bool trunc)
{
    if (count <= 0)
        return hipSuccess;
    // a truncated transform longer than one sub-transform is split in four
    const int logM = trunc && logN > NttLimits<W>::max_log_m
                         ? logN - 2
                         : (logN < NttLimits<W>::max_log_m ? logN : NttLimits<W>::max_log_m);
    const int logR = logN - logM;
    if (logR > NttLimits<W>::max_log_r || logP > logN)
        return hipErrorInvalidValue;
    if (trunc && (logR > 2 || logN < kTruncMinLog || (logR > 0 && logM < kRegMinLog)))
        return hipErrorInvalidValue;
    const int64_t units = count << logR;
    if (logR > 2) {
        // long transform: column pass into tmp, then the rows
        if (!tmp || !ct)
            return hipErrorInvalidValue;
        hipError_t e = launch_cols_fwd<Win, W>(in, in_stride, in_len, tmp, count, logM, logR, *ct,
                                               p, s);
        if (e != hipSuccess)
            return e;
        return dispatch_fwd_reg<Win, W>(logM, tune.fwd_log_e, tmp, (int64_t)1 << logN,
                                        (int64_t)1 << logN, out, units, logR, logP, tab, p, s, 1);
    }
    if (logM >= kRegMinLog)
        return dispatch_fwd_reg<Win, W>(logM, tune.fwd_log_e, in, in_stride, in_len, out, units,
                                        logR, logP, tab_for_sub(tab, logN, logM), p, s,
                                        trunc ? 2 : 0);
    const int lognb = pick_lognb(logM);
    const int64_t blocks = (units + (1 << lognb) - 1) >> lognb;
    const size_t lds = ((size_t)1 << (logM + lognb)) * sizeof(W);
    hipLaunchKernelGGL((ntt_fwd_kernel<Win, W>), dim3((unsigned)blocks), dim3(kThreads), lds, s, in,
                       in_stride, in_len, out, units, logN, logR, lognb, logP, tab, p, trunc ? 1 : 0);
    return hipGetLastError();
}

template <typename W, typename Wout>
hipError_t launch_ntt_inv(const W *in, Wout *out, int64_t out_stride, int64_t lo, int64_t hi,
                          int64_t count, int logN, int logP, const NttTune &tune,
                          const NttTab<W> &tab, W p, W *tmp, hipStream_t s, const ColTab<W> *ct,
                          bool trunc)
{
    const int64_t N = (int64_t)1 << logN;
    if (hi > N)
        hi = N;
    if (count <= 0 || hi <= lo)
        return hipSuccess;
    const int logM = trunc && logN > NttLimits<W>::max_log_m
                         ? logN - 2
                         : (logN < NttLimits<W>::max_log_m ? logN : NttLimits<W>::max_log_m);
    const int logR = logN - logM;
    if (logR > NttLimits<W>::max_log_r || (logR > 0 && tmp == nullptr) || logP > logN ||
        (logR > 2 && (!ct || logM < kRegMinLog)) ||
        (trunc && (logR > 2 || logN < kTruncMinLog || hi > 3 * (N >> 2) ||
                   (logR > 0 && logM < kRegMinLog))))
        return hipErrorInvalidValue;
    const int64_t units = count << logR;
    hipError_t err;
    if (logM >= kRegMinLog) {
        err = dispatch_inv_reg<W, Wout>(logM, tune.inv_log_e, in, out, out_stride, lo, hi, tmp,
                                        units, logR, logP, tab_for_sub(tab, logN, logM), p, s,
                                        trunc ? 1 : 0);
    } else {
        const int lognb = pick_lognb(logM);
        const int64_t blocks = (units + (1 << lognb) - 1) >> lognb;
        const size_t lds = ((size_t)1 << (logM + lognb)) * sizeof(W);
        hipLaunchKernelGGL((ntt_inv_kernel<W, Wout>), dim3((unsigned)blocks), dim3(kThreads), lds,
                           s, in, out, out_stride, lo, hi, tmp, units, logN, logR, lognb, logP,
                           tab, p, trunc ? 1 : 0);
        err = hipGetLastError();
    }
    if (err != hipSuccess || logR == 0)
        return err;
    if (logR > 2)
        return launch_cols_inv<W, Wout>(tmp, out, out_stride, lo, hi, count, logM, logR, *ct,
                                        tab.ninv, tab.ninvq, p, s);
    if (trunc) {
        int64_t tblocks = ((count << (logN - 2)) + kThreads - 1) / kThreads;
        if (tblocks > 65536)
            tblocks = 65536;
        hipLaunchKernelGGL((ntt_merge_trunc_kernel<W, Wout>), dim3((unsigned)tblocks), dim3(kThreads),
                           0, s, tmp, out, out_stride, lo, hi, count, logN, tab, p);
        return hipGetLastError();
    }
    const int64_t total = count * (hi - lo);
    int64_t mblocks = (total + kThreads - 1) / kThreads;
    if (mblocks > 65536)
        mblocks = 65536;
    hipLaunchKernelGGL((ntt_merge_kernel<W, Wout>), dim3((unsigned)mblocks), dim3(kThreads), 0, s,
                       tmp, out, out_stride, lo, hi, count, logN, logR, tab, p);
    return hipGetLastError();
}

// ===========================================================================
// Segment-table launches (batched products): same workgroup bodies, the
// operand matrix of each workgroup looked up by a binary search on block0.
// Blocks go through the same XCD-aware order as the direct launches (when
// the grid is a multiple of 8), so that neighbouring entries, which share
// cache lines in the blocked layout, are handled on one XCD.
// ===========================================================================
namespace {

template <typename S>
__device__ __forceinline__ const S &seg_of(const S *segs, int nseg, int64_t b)
{
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].block0 <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return segs[lo];
}

template <typename S>
__device__ __forceinline__ const S &mseg_of(const S *segs, int nseg, int64_t b)
{
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (segs[mid].mblock0 <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return segs[lo];
}

template <typename Win, typename W>
__global__ void __launch_bounds__(kThreads)
ntt_fwd_segs_lds(const FwdSeg<Win, W> *__restrict__ segs, int nseg, int logN, int logR, int lognb,
                 int logP, NttTab<W> tab, W p)
{
    const int64_t b = xcd_unit(blockIdx.x, gridDim.x);
    const FwdSeg<Win, W> &sg = seg_of(segs, nseg, b);
    fwd_lds_body<Win, W>(sg.in, sg.in_stride, sg.in_len, sg.out, sg.count << logR,
                         (b - sg.block0) << lognb, logN, logR, lognb, logP, tab, p, sg.trunc != 0);
}

template <typename W, typename Wout>
__global__ void __launch_bounds__(kThreads)
ntt_inv_segs_lds(const InvSeg<W, Wout> *__restrict__ segs, int nseg, int logN, int logR,
                 int lognb, int logP, NttTab<W> tab, W p)
{
    const int64_t b = xcd_unit(blockIdx.x, gridDim.x);
    const InvSeg<W, Wout> &sg = seg_of(segs, nseg, b);
    inv_lds_body<W, Wout>(sg.in, sg.out, sg.out_stride, sg.lo, sg.hi, sg.tmp, sg.count << logR,
                          (b - sg.block0) << lognb, logN, logR, lognb, logP, tab, p, sg.trunc != 0);
}

// every segment of one launch has the launch's TRUNC (see ntt_fwd_reg)
template <typename Win, typename W, bool LOOSE, int L, int LOGE, bool TRUNC>
__global__ void __launch_bounds__(1 << (L - LOGE)) __attribute__((amdgpu_waves_per_eu(LOGE == 5 ? kFwdWaves32 : 1)))
ntt_fwd_segs_reg(const FwdSeg<Win, W> *__restrict__ segs, int nseg, int logR, int logP,
                 NttTab<W> tab, W p)
{
    const int64_t b = xcd_unit(blockIdx.x, gridDim.x);
    const FwdSeg<Win, W> &sg = seg_of(segs, nseg, b);
    fwd_reg_body<Win, W, LOOSE, L, LOGE>(sg.in, sg.in_stride, sg.in_len, sg.out, sg.count << logR,
                                         b - sg.block0, logR, logP, tab, p, false, TRUNC);
}

template <typename W, typename Wout, bool LOOSE, int L, int LOGE>
__global__ void __launch_bounds__(1 << (L - LOGE)) __attribute__((amdgpu_waves_per_eu(LOGE == 4 ? kInvWaves : 1)))
ntt_inv_segs_reg(const InvSeg<W, Wout> *__restrict__ segs, int nseg, int logR, int logP,
                 NttTab<W> tab, W p)
{
    const int64_t b = xcd_unit(blockIdx.x, gridDim.x);
    const InvSeg<W, Wout> &sg = seg_of(segs, nseg, b);
    inv_reg_body<W, Wout, LOOSE, L, LOGE>(sg.in, sg.out, sg.out_stride, sg.lo, sg.hi, sg.tmp,
                                          sg.count << logR, b - sg.block0, logR, logP, tab, p,
                                          sg.trunc != 0);
}

template <typename W, typename Wout>
__global__ void __launch_bounds__(kThreads)
ntt_merge_segs(const InvSeg<W, Wout> *__restrict__ segs, int nseg, int logN, int logR,
               NttTab<W> tab, W p)
{
    const InvSeg<W, Wout> &sg = mseg_of(segs, nseg, blockIdx.x);
    const int64_t idx = ((int64_t)blockIdx.x - sg.mblock0) * kThreads + threadIdx.x;
    if (sg.trunc) {
        if (idx < (sg.count << (logN - 2)))
            merge_trunc_one<W, Wout>(sg.tmp, sg.out, sg.out_stride, sg.lo, sg.hi, idx, logN, tab,
                                     p);
        return;
    }
    const int64_t width = sg.hi - sg.lo;
    if (idx < sg.count * width)
        merge_one<W, Wout>(sg.tmp, sg.out, sg.out_stride, sg.lo, width, idx, logN, logR, tab, p);
}

template <typename Kern>
hipError_t set_lds_attr(Kern *k, size_t lds, bool &done)
{
    if (done || lds <= 64 * 1024)
        return hipSuccess;
    hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e == hipSuccess)
        done = true;
    return e;
}

template <typename Win, typename W, bool LOOSE, int L, int LOGE, bool TRUNC>
hipError_t fwd_segs_reg4(const FwdSeg<Win, W> *segs, int nseg, int64_t blocks, int logR, int logP,
                         const NttTab<W> &tab, W p, hipStream_t s)
{
    const size_t lds = (size_t)((1 << L) + (1 << L) / 32) * sizeof(W);
    static bool attr = false;
    hipError_t e = set_lds_attr(ntt_fwd_segs_reg<Win, W, LOOSE, L, LOGE, TRUNC>, lds, attr);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL((ntt_fwd_segs_reg<Win, W, LOOSE, L, LOGE, TRUNC>), dim3((unsigned)blocks),
                       dim3(1 << (L - LOGE)), lds, s, segs, nseg, logR, logP, tab, p);
    return hipGetLastError();
}

template <typename Win, typename W, bool LOOSE, int L, int LOGE>
hipError_t fwd_segs_reg3(const FwdSeg<Win, W> *segs, int nseg, int64_t blocks, int logR, int logP,
                         const NttTab<W> &tab, W p, hipStream_t s, bool trunc)
{
    if (trunc)
        return fwd_segs_reg4<Win, W, LOOSE, L, LOGE, true>(segs, nseg, blocks, logR, logP, tab, p,
                                                           s);
    return fwd_segs_reg4<Win, W, LOOSE, L, LOGE, false>(segs, nseg, blocks, logR, logP, tab, p, s);
}

template <typename W, typename Wout, bool LOOSE, int L, int LOGE>
hipError_t inv_segs_reg3(const InvSeg<W, Wout> *segs, int nseg, int64_t blocks, int logR, int logP,
                         const NttTab<W> &tab, W p, hipStream_t s)
{
    const size_t lds = (size_t)((1 << L) + (1 << L) / 32) * sizeof(W);
    static bool attr = false;
    hipError_t e = set_lds_attr(ntt_inv_segs_reg<W, Wout, LOOSE, L, LOGE>, lds, attr);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL((ntt_inv_segs_reg<W, Wout, LOOSE, L, LOGE>), dim3((unsigned)blocks),
                       dim3(1 << (L - LOGE)), lds, s, segs, nseg, logR, logP, tab, p);
    return hipGetLastError();
}

template <typename Win, typename W, bool LOOSE, int L>
hipError_t fwd_segs_reg2(int logE, const FwdSeg<Win, W> *segs, int nseg, int64_t blocks, int logR,
                         int logP, const NttTab<W> &tab, W p, hipStream_t s, bool trunc)
{
    if (logE == 5 && L >= 12)
        return fwd_segs_reg3<Win, W, LOOSE, L, (L >= 12 ? 5 : 4)>(segs, nseg, blocks, logR, logP,
                                                                 tab, p, s, trunc);
    return fwd_segs_reg3<Win, W, LOOSE, L, 4>(segs, nseg, blocks, logR, logP, tab, p, s, trunc);
}

template <typename W, typename Wout, bool LOOSE, int L>
hipError_t inv_segs_reg2(int logE, const InvSeg<W, Wout> *segs, int nseg, int64_t blocks,
                         int logR, int logP, const NttTab<W> &tab, W p, hipStream_t s)
{
    if (logE == 5 && L >= 12)
        return inv_segs_reg3<W, Wout, LOOSE, L, (L >= 12 ? 5 : 4)>(segs, nseg, blocks, logR, logP,
                                                                  tab, p, s);
    return inv_segs_reg3<W, Wout, LOOSE, L, 4>(segs, nseg, blocks, logR, logP, tab, p, s);
}

template <typename Win, typename W, int L>
hipError_t fwd_segs_reg(int logE, const FwdSeg<Win, W> *segs, int nseg, int64_t blocks, int logR,
                        int logP, const NttTab<W> &tab, W p, hipStream_t s, bool trunc)
{
    if constexpr (sizeof(W) == 8) {
        if ((uint64_t)p < ((uint64_t)1 << 60))
            return fwd_segs_reg2<Win, W, true, L>(logE, segs, nseg, blocks, logR, logP, tab, p, s,
                                                  trunc);
    }
    return fwd_segs_reg2<Win, W, false, L>(logE, segs, nseg, blocks, logR, logP, tab, p, s, trunc);
}

template <typename W, typename Wout, int L>
hipError_t inv_segs_reg(int logE, const InvSeg<W, Wout> *segs, int nseg, int64_t blocks, int logR,
                        int logP, const NttTab<W> &tab, W p, hipStream_t s)
{
    if constexpr (sizeof(W) == 8) {
        if ((uint64_t)p < ((uint64_t)1 << 60))
            return inv_segs_reg2<W, Wout, true, L>(logE, segs, nseg, blocks, logR, logP, tab, p, s);
    }
    return inv_segs_reg2<W, Wout, false, L>(logE, segs, nseg, blocks, logR, logP, tab, p, s);
}

template <typename W>
int seg_log_m(int logN, bool trunc)
{
    // a truncated product longer than one sub-transform is split in four
    if (trunc && logN > NttLimits<W>::max_log_m)
        return logN - 2;
    return logN < NttLimits<W>::max_log_m ? logN : NttLimits<W>::max_log_m;
}

}  // namespace

template <typename W>
int64_t ntt_seg_blocks(int64_t count, int logN, bool trunc)
{
    const int logM = seg_log_m<W>(logN, trunc);
    const int64_t units = count << (logN - logM);
    if (logM >= kRegMinLog)
        return units;
    const int lognb = pick_lognb(logM);
    return (units + (1 << lognb) - 1) >> lognb;
}

template <typename W>
int64_t merge_seg_blocks(int64_t count, int64_t width, int logN, bool trunc)
{
    if (logN <= NttLimits<W>::max_log_m || width <= 0)
        return 0;
    if (trunc)  // one item per quadruple of outputs
        return ((count << (logN - 2)) + kThreads - 1) / kThreads;
    return (count * width + kThreads - 1) / kThreads;
}

template <typename Win, typename W>
hipError_t launch_ntt_fwd_segs(const FwdSeg<Win, W> *segs, int nseg, int64_t blocks, int logN,
                               int logP, int log_e, const NttTab<W> &tab, W p, hipStream_t s,
                               bool trunc)
{
    if (nseg <= 0 || blocks <= 0)
        return hipSuccess;
    const int logM = seg_log_m<W>(logN, trunc);
    const int logR = logN - logM;
    if (logR > 2 || logP > logN || blocks > INT32_MAX)
        return hipErrorInvalidValue;
    const NttTab<W> st = tab_for_sub(tab, logN, logM);
    if (logM >= kRegMinLog) {
        switch (logM) {
        case 11:
            return fwd_segs_reg<Win, W, 11>(log_e, segs, nseg, blocks, logR, logP, st, p, s,
                                             trunc);
        case 12:
            return fwd_segs_reg<Win, W, 12>(log_e, segs, nseg, blocks, logR, logP, st, p, s,
                                             trunc);
        case 13:
            return fwd_segs_reg<Win, W, 13>(log_e, segs, nseg, blocks, logR, logP, st, p, s,
                                             trunc);
        case 14:
            if constexpr (sizeof(W) == 4)
                return fwd_segs_reg<Win, W, 14>(log_e, segs, nseg, blocks, logR, logP, st, p, s,
                                                 trunc);
            [[fallthrough]];
        default:
            return hipErrorInvalidValue;
        }
    }
    const int lognb = pick_lognb(logM);
    const size_t lds = ((size_t)1 << (logM + lognb)) * sizeof(W);
    hipLaunchKernelGGL((ntt_fwd_segs_lds<Win, W>), dim3((unsigned)blocks), dim3(kThreads), lds, s,
                       segs, nseg, logN, logR, lognb, logP, tab, p);
    return hipGetLastError();
}

template <typename W, typename Wout>
hipError_t launch_ntt_inv_segs(const InvSeg<W, Wout> *segs, int nseg, int64_t blocks,
                               int64_t mblocks, int logN, int logP, int log_e,
                               const NttTab<W> &tab, W p, hipStream_t s, bool trunc)
{
    if (nseg <= 0 || blocks <= 0)
        return hipSuccess;
    const int logM = seg_log_m<W>(logN, trunc);
    const int logR = logN - logM;
    if (logR > 2 || logP > logN || blocks > INT32_MAX || mblocks > INT32_MAX)
        return hipErrorInvalidValue;
    const NttTab<W> st = tab_for_sub(tab, logN, logM);
    hipError_t err;
    if (logM >= kRegMinLog) {
        switch (logM) {
        case 11:
            err = inv_segs_reg<W, Wout, 11>(log_e, segs, nseg, blocks, logR, logP, st, p, s);
            break;
        case 12:
            err = inv_segs_reg<W, Wout, 12>(log_e, segs, nseg, blocks, logR, logP, st, p, s);
            break;
        case 13:
            err = inv_segs_reg<W, Wout, 13>(log_e, segs, nseg, blocks, logR, logP, st, p, s);
            break;
        case 14:
            if constexpr (sizeof(W) == 4) {
                err = inv_segs_reg<W, Wout, 14>(log_e, segs, nseg, blocks, logR, logP, st, p, s);
                break;
            }
            [[fallthrough]];
        default:
            return hipErrorInvalidValue;
        }
    } else {
        const int lognb = pick_lognb(logM);
        const size_t lds = ((size_t)1 << (logM + lognb)) * sizeof(W);
        hipLaunchKernelGGL((ntt_inv_segs_lds<W, Wout>), dim3((unsigned)blocks), dim3(kThreads), lds,
                           s, segs, nseg, logN, logR, lognb, logP, tab, p);
        err = hipGetLastError();
    }
    if (err != hipSuccess || logR == 0 || mblocks <= 0)
        return err;
    hipLaunchKernelGGL((ntt_merge_segs<W, Wout>), dim3((unsigned)mblocks), dim3(kThreads), 0, s,
                       segs, nseg, logN, logR, tab, p);
    return hipGetLastError();
}

template int64_t ntt_seg_blocks<uint64_t>(int64_t, int, bool);
template int64_t ntt_seg_blocks<uint32_t>(int64_t, int, bool);
template int64_t merge_seg_blocks<uint64_t>(int64_t, int64_t, int, bool);
template int64_t merge_seg_blocks<uint32_t>(int64_t, int64_t, int, bool);
template hipError_t launch_ntt_fwd_segs<uint64_t, uint64_t>(const FwdSeg<uint64_t, uint64_t> *, int,
                                                            int64_t, int, int, int,
                                                            const NttTab<uint64_t> &, uint64_t,
                                                            hipStream_t, bool);
template hipError_t launch_ntt_fwd_segs<uint64_t, uint32_t>(const FwdSeg<uint64_t, uint32_t> *, int,
                                                            int64_t, int, int, int,
                                                            const NttTab<uint32_t> &, uint32_t,
                                                            hipStream_t, bool);
template hipError_t launch_ntt_inv_segs<uint64_t, uint64_t>(const InvSeg<uint64_t, uint64_t> *, int,
                                                            int64_t, int64_t, int, int, int,
                                                            const NttTab<uint64_t> &, uint64_t,
                                                            hipStream_t, bool);
template hipError_t launch_ntt_inv_segs<uint32_t, uint64_t>(const InvSeg<uint32_t, uint64_t> *, int,
                                                            int64_t, int64_t, int, int, int,
                                                            const NttTab<uint32_t> &, uint32_t,
                                                            hipStream_t, bool);

template hipError_t launch_ntt_fwd<uint64_t, uint64_t>(const uint64_t *, int64_t, int64_t,
                                                       uint64_t *, int64_t, int, int,
                                                       const NttTune &, const NttTab<uint64_t> &,
                                                       uint64_t, hipStream_t, uint64_t *,
                                                       const ColTab<uint64_t> *, bool);
template hipError_t launch_ntt_fwd<uint64_t, uint32_t>(const uint64_t *, int64_t, int64_t,
                                                       uint32_t *, int64_t, int, int,
                                                       const NttTune &, const NttTab<uint32_t> &,
                                                       uint32_t, hipStream_t, uint64_t *,
                                                       const ColTab<uint32_t> *, bool);
template hipError_t launch_ntt_inv<uint64_t, uint64_t>(const uint64_t *, uint64_t *, int64_t,
                                                       int64_t, int64_t, int64_t, int, int,
                                                       const NttTune &, const NttTab<uint64_t> &,
                                                       uint64_t, uint64_t *, hipStream_t,
                                                       const ColTab<uint64_t> *, bool);
template hipError_t launch_ntt_inv<uint32_t, uint64_t>(const uint32_t *, uint64_t *, int64_t,
                                                       int64_t, int64_t, int64_t, int, int,
                                                       const NttTune &, const NttTab<uint32_t> &,
                                                       uint32_t, uint32_t *, hipStream_t,
                                                       const ColTab<uint32_t> *, bool);

}  // namespace pmlg
