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

template <typename W>
__device__ __forceinline__ W tw_mul(W x, const W *w, const W *wq, int64_t idx, W p)
{
    return shoup_lazy(x, w[idx], wq[idx], p);
}

template <typename Win, typename W>
__global__ void __launch_bounds__(kThreads)
ntt_fwd_kernel(const Win *__restrict__ in, int64_t in_stride, int64_t in_len,
               W *__restrict__ out, int64_t units, int logN, int logR, int lognb,
               NttTab<W> tab, W p)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    W *s = reinterpret_cast<W *>(smem_raw);
    const int logM = logN - logR;
    const int M = 1 << logM;
    const int R = 1 << logR;
    const int nb = 1 << lognb;
    const int64_t N = (int64_t)1 << logN;
    const W p2 = p + p;
    const int64_t u0 = (int64_t)blockIdx.x * nb;

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

    // store: unit u = (e, g) -> out[e*N + g*M + i]
    for (int idx = threadIdx.x; idx < nb * M; idx += kThreads) {
        const int t = idx >> logM;
        const int i1 = idx & (M - 1);
        const int64_t u = u0 + t;
        if (u < units) {
            const int64_t e = u >> logR;
            const int g = (int)(u & (R - 1));
            out[e * N + (int64_t)g * M + i1] = reduce_2p<W>(s[idx], p);
        }
    }
}

template <typename W, typename Wout>
__global__ void __launch_bounds__(kThreads)
ntt_inv_kernel(const W *__restrict__ in, Wout *__restrict__ out, int64_t out_stride, int64_t lo,
               int64_t hi, W *__restrict__ tmp, int64_t units, int logN, int logR, int lognb,
               NttTab<W> tab, W p)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    W *s = reinterpret_cast<W *>(smem_raw);
    const int logM = logN - logR;
    const int M = 1 << logM;
    const int R = 1 << logR;
    const int nb = 1 << lognb;
    const int64_t N = (int64_t)1 << logN;
    const W p2 = p + p;
    const int64_t u0 = (int64_t)blockIdx.x * nb;

    for (int idx = threadIdx.x; idx < nb * M; idx += kThreads) {
        const int t = idx >> logM;
        const int i1 = idx & (M - 1);
        const int64_t u = u0 + t;
        W v = 0;
        if (u < units) {
            const int64_t e = u >> logR;
            const int g = (int)(u & (R - 1));
            v = in[e * N + (int64_t)g * M + i1];
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
                    (Wout)shoup<W>(s[idx], tab.ninv, tab.ninvq, p);
        } else {
            const int g = (int)(u & (R - 1));
            tmp[e * N + (int64_t)g * M + i1] = s[idx];
        }
    }
}

// x[o] = N^-1 * sum_g w^(-(o g) mod N) z_g[o mod M], o in [lo, hi)
template <typename W, typename Wout>
__global__ void __launch_bounds__(kThreads)
ntt_merge_kernel(const W *__restrict__ z, Wout *__restrict__ out, int64_t out_stride, int64_t lo,
                 int64_t hi, int64_t count, int logN, int logR, NttTab<W> tab, W p)
{
    const int64_t N = (int64_t)1 << logN;
    const int logM = logN - logR;
    const int64_t M = (int64_t)1 << logM;
    const int R = 1 << logR;
    const int64_t width = hi - lo;
    const int64_t total = count * width;
    const W p2 = p + p;
    for (int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * kThreads) {
        const int64_t e = idx / width;
        const int64_t o = lo + (idx - e * width);
        const int64_t i1 = o & (M - 1);
        const W *ze = z + e * N;
        W acc = reduce_2p<W>(ze[i1], p);
        for (int g = 1; g < R; ++g) {
            const int64_t widx = (o * g) & (N - 1);
            acc = add_2p<W>(acc, tw_mul<W>(ze[(int64_t)g * M + i1], tab.itw, tab.itwq, widx, p), p2);
        }
        out[e * out_stride + (o - lo)] = (Wout)shoup<W>(acc, tab.ninv, tab.ninvq, p);
    }
}

inline int pick_lognb(int logM)
{
    const int want = 11;  // at least 2048 elements per workgroup
    return logM >= want ? 0 : want - logM;
}

}  // namespace

template <typename Win, typename W>
hipError_t launch_ntt_fwd(const Win *in, int64_t in_stride, int64_t in_len, W *out,
                          int64_t count, int logN, const NttTab<W> &tab, W p, hipStream_t s)
{
    if (count <= 0)
        return hipSuccess;
    const int logM = logN < NttLimits<W>::max_log_m ? logN : NttLimits<W>::max_log_m;
    const int logR = logN - logM;
    if (logR > 2)
        return hipErrorInvalidValue;
    const int lognb = pick_lognb(logM);
    const int64_t units = count << logR;
    const int64_t blocks = (units + (1 << lognb) - 1) >> lognb;
    const size_t lds = ((size_t)1 << (logM + lognb)) * sizeof(W);
    hipLaunchKernelGGL((ntt_fwd_kernel<Win, W>), dim3((unsigned)blocks), dim3(kThreads), lds, s, in,
                       in_stride, in_len, out, units, logN, logR, lognb, tab, p);
    return hipGetLastError();
}

template <typename W, typename Wout>
hipError_t launch_ntt_inv(const W *in, Wout *out, int64_t out_stride, int64_t lo, int64_t hi,
                          int64_t count, int logN, const NttTab<W> &tab, W p, W *tmp,
                          hipStream_t s)
{
    const int64_t N = (int64_t)1 << logN;
    if (hi > N)
        hi = N;
    if (count <= 0 || hi <= lo)
        return hipSuccess;
    const int logM = logN < NttLimits<W>::max_log_m ? logN : NttLimits<W>::max_log_m;
    const int logR = logN - logM;
    if (logR > 2 || (logR > 0 && tmp == nullptr))
        return hipErrorInvalidValue;
    const int lognb = pick_lognb(logM);
    const int64_t units = count << logR;
    const int64_t blocks = (units + (1 << lognb) - 1) >> lognb;
    const size_t lds = ((size_t)1 << (logM + lognb)) * sizeof(W);
    hipLaunchKernelGGL((ntt_inv_kernel<W, Wout>), dim3((unsigned)blocks), dim3(kThreads), lds, s,
                       in, out, out_stride, lo, hi, tmp, units, logN, logR, lognb, tab, p);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess || logR == 0)
        return err;
    const int64_t total = count * (hi - lo);
    int64_t mblocks = (total + kThreads - 1) / kThreads;
    if (mblocks > 65536)
        mblocks = 65536;
    hipLaunchKernelGGL((ntt_merge_kernel<W, Wout>), dim3((unsigned)mblocks), dim3(kThreads), 0, s,
                       tmp, out, out_stride, lo, hi, count, logN, logR, tab, p);
    return hipGetLastError();
}

template hipError_t launch_ntt_fwd<uint64_t, uint64_t>(const uint64_t *, int64_t, int64_t,
                                                       uint64_t *, int64_t, int,
                                                       const NttTab<uint64_t> &, uint64_t,
                                                       hipStream_t);
template hipError_t launch_ntt_fwd<uint64_t, uint32_t>(const uint64_t *, int64_t, int64_t,
                                                       uint32_t *, int64_t, int,
                                                       const NttTab<uint32_t> &, uint32_t,
                                                       hipStream_t);
template hipError_t launch_ntt_inv<uint64_t, uint64_t>(const uint64_t *, uint64_t *, int64_t,
                                                       int64_t, int64_t, int64_t, int,
                                                       const NttTab<uint64_t> &, uint64_t,
                                                       uint64_t *, hipStream_t);
template hipError_t launch_ntt_inv<uint32_t, uint64_t>(const uint32_t *, uint64_t *, int64_t,
                                                       int64_t, int64_t, int64_t, int,
                                                       const NttTab<uint32_t> &, uint32_t,
                                                       uint32_t *, hipStream_t);

}  // namespace pmlg
