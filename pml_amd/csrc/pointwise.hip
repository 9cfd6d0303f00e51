// Pointwise modular matrix products at every evaluation point:
//   C_hat[i][j][q] = sum_k A_hat[i][k][q] * B_hat[k][j][q] mod p.
//
// Two kernels:
//  * plain: each lane owns one point q and a TM x TN block of outputs; the
//    dot products are accumulated exactly in 128 bits (64-bit words) or
//    64 bits (32-bit words) and reduced once per 16 terms;
//  * mfma: residues are split into L balanced signed 8-bit limbs
//    (x = sum_t l_t 2^(8t), l_t in [-128,127]); for every weight w the
//    int8 MFMA chain T_w = sum_{s+t=w} A_s * B_t accumulates exactly in int32
//    (|T_w| <= 32 * L * 2^14 per 32-deep K chunk), the T_w are folded into
//    three (or two) signed 64-bit partial sums by shifts, and the result is
//    reduced mod p with Barrett/Shoup steps.  No floating point is involved.
#include "internal.hpp"

namespace pmlg {

namespace {

constexpr int kThreads = 256;
constexpr int kLdsBudget = 128 * 1024;  // bytes of LDS for limb planes

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// plain kernel
// ---------------------------------------------------------------------------
constexpr int TM = 2, TN = 4;

__device__ __forceinline__ void plain_fold(uint64_t &hi, uint64_t &lo, const ModConst &mc)
{
    lo = reduce128(hi, lo, mc);
    hi = 0;
}

__global__ void __launch_bounds__(kThreads)
pw_plain_u64(const uint64_t *__restrict__ Ah, const uint64_t *__restrict__ Bh,
             uint64_t *__restrict__ Ch, int m, int k, int n, int64_t N, ModConst mc)
{
    const int64_t q = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int tiles_n = (n + TN - 1) / TN;
    const int tile = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i0 = (tile / tiles_n) * TM;
    const int j0 = (tile % tiles_n) * TN;
    if (q >= N || i0 >= m)
        return;
    uint64_t hi[TM][TN], lo[TM][TN];
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c)
            hi[r][c] = lo[r][c] = 0;
    int cnt = 0;
    for (int kk = 0; kk < k; ++kk) {
        uint64_t a[TM], b[TN];
#pragma unroll
        for (int r = 0; r < TM; ++r)
            a[r] = (i0 + r < m) ? Ah[((int64_t)(i0 + r) * k + kk) * N + q] : 0;
#pragma unroll
        for (int c = 0; c < TN; ++c)
            b[c] = (j0 + c < n) ? Bh[((int64_t)kk * n + j0 + c) * N + q] : 0;
#pragma unroll
        for (int r = 0; r < TM; ++r)
#pragma unroll
            for (int c = 0; c < TN; ++c)
                mac128(hi[r][c], lo[r][c], a[r], b[c]);
        if (++cnt == 15) {  // 15 products < 2^124 each plus a residue < p: no overflow
#pragma unroll
            for (int r = 0; r < TM; ++r)
#pragma unroll
                for (int c = 0; c < TN; ++c)
                    plain_fold(hi[r][c], lo[r][c], mc);
            cnt = 0;
        }
    }
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c)
            if (i0 + r < m && j0 + c < n)
                Ch[((int64_t)(i0 + r) * n + j0 + c) * N + q] = reduce128(hi[r][c], lo[r][c], mc);
}

__global__ void __launch_bounds__(kThreads)
pw_plain_u32(const uint32_t *__restrict__ Ah, const uint32_t *__restrict__ Bh,
             uint32_t *__restrict__ Ch, int m, int k, int n, int64_t N, ModConst mc)
{
    const int64_t q = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int tiles_n = (n + TN - 1) / TN;
    const int tile = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int i0 = (tile / tiles_n) * TM;
    const int j0 = (tile % tiles_n) * TN;
    if (q >= N || i0 >= m)
        return;
    uint64_t acc[TM][TN];
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c)
            acc[r][c] = 0;
    int cnt = 0;
    for (int kk = 0; kk < k; ++kk) {
        uint64_t a[TM], b[TN];
#pragma unroll
        for (int r = 0; r < TM; ++r)
            a[r] = (i0 + r < m) ? Ah[((int64_t)(i0 + r) * k + kk) * N + q] : 0;
#pragma unroll
        for (int c = 0; c < TN; ++c)
            b[c] = (j0 + c < n) ? Bh[((int64_t)kk * n + j0 + c) * N + q] : 0;
#pragma unroll
        for (int r = 0; r < TM; ++r)
#pragma unroll
            for (int c = 0; c < TN; ++c)
                acc[r][c] += a[r] * b[c];  // < 2^60 each
        if (++cnt == 15) {
#pragma unroll
            for (int r = 0; r < TM; ++r)
#pragma unroll
                for (int c = 0; c < TN; ++c)
                    acc[r][c] = barrett(acc[r][c], mc.bm, mc.p);
            cnt = 0;
        }
    }
#pragma unroll
    for (int r = 0; r < TM; ++r)
#pragma unroll
        for (int c = 0; c < TN; ++c)
            if (i0 + r < m && j0 + c < n)
                Ch[((int64_t)(i0 + r) * n + j0 + c) * N + q] =
                    (uint32_t)barrett(acc[r][c], mc.bm, mc.p);
}

// ---------------------------------------------------------------------------
// int8-limb MFMA kernel
// ---------------------------------------------------------------------------

// Balanced signed base-256 digits of x (x < 2^61 for L = 8, x < 2^30 for L = 4).
template <int L, typename W>
__device__ __forceinline__ void split_limbs(W x, int (&l)[L])
{
    int carry = 0;
#pragma unroll
    for (int t = 0; t < L; ++t) {
        int v = (int)((x >> (8 * t)) & 255) + carry;
        carry = (v + 128) >> 8;
        l[t] = v - (carry << 8);
    }
}

template <typename W, int PB>
__device__ __forceinline__ void load_run(const W *__restrict__ src, W (&v)[PB])
{
    if constexpr (PB * sizeof(W) >= 16) {
        const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
        constexpr int n4 = PB * sizeof(W) / 16;
#pragma unroll
        for (int t = 0; t < n4; ++t) {
            uint4 x = s4[t];
            W *d = v + t * (16 / sizeof(W));
            if constexpr (sizeof(W) == 8) {
                d[0] = ((uint64_t)x.y << 32) | x.x;
                d[1] = ((uint64_t)x.w << 32) | x.z;
            } else {
                d[0] = x.x;
                d[1] = x.y;
                d[2] = x.z;
                d[3] = x.w;
            }
        }
    } else {
#pragma unroll
        for (int t = 0; t < PB; ++t)
            v[t] = src[t];
    }
}

// Output layout of v_mfma_i32_32x32x32_i8: register r of lane l holds
// D[row][col] with col = l % 32 and row = 8 (r / 4) + 4 (l / 32) + r % 4.
__device__ __forceinline__ int mfma_row(int r, int lane)
{
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

template <typename W, int L, int PB>
__global__ void __launch_bounds__(kThreads)
pw_mfma_kernel(const W *__restrict__ Ah, const W *__restrict__ Bh, W *__restrict__ Ch, int m,
               int k, int n, int64_t N, ModConst mc)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const int Mp = (m + 31) & ~31, Kp = (k + 31) & ~31, Np = (n + 31) & ~31;
    const int Kq = Kp >> 2;
    unsigned char *la = lds;                           // [PB][L][Mp][Kp]
    unsigned char *lb = lds + (size_t)PB * L * Mp * Kp;  // [PB][L][Np][Kp]  (B transposed)
    const int64_t q0 = (int64_t)blockIdx.x * PB;
    const int tid = threadIdx.x;

    // ---- A limbs: one thread per (row i, 4 consecutive k) ----
    for (int quad = tid; quad < Mp * Kq; quad += kThreads) {
        const int i = quad / Kq;
        const int kq = quad - i * Kq;
        uint32_t words[PB][L];
#pragma unroll
        for (int pt = 0; pt < PB; ++pt)
#pragma unroll
            for (int t = 0; t < L; ++t)
                words[pt][t] = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kk = 4 * kq + u;
            if (i < m && kk < k) {
                W v[PB];
                load_run<W, PB>(Ah + ((int64_t)i * k + kk) * N + q0, v);
#pragma unroll
                for (int pt = 0; pt < PB; ++pt) {
                    int l[L];
                    split_limbs<L, W>(v[pt], l);
#pragma unroll
                    for (int t = 0; t < L; ++t)
                        words[pt][t] |= (uint32_t)(l[t] & 255) << (8 * u);
                }
            }
        }
#pragma unroll
        for (int pt = 0; pt < PB; ++pt)
#pragma unroll
            for (int t = 0; t < L; ++t)
                *reinterpret_cast<uint32_t *>(la + ((size_t)(pt * L + t) * Mp + i) * Kp + 4 * kq) =
                    words[pt][t];
    }
    // ---- B limbs, transposed: one thread per (column j, 4 consecutive k) ----
    for (int quad = tid; quad < Np * Kq; quad += kThreads) {
        const int j = quad / Kq;
        const int kq = quad - j * Kq;
        uint32_t words[PB][L];
#pragma unroll
        for (int pt = 0; pt < PB; ++pt)
#pragma unroll
            for (int t = 0; t < L; ++t)
                words[pt][t] = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kk = 4 * kq + u;
            if (j < n && kk < k) {
                W v[PB];
                load_run<W, PB>(Bh + ((int64_t)kk * n + j) * N + q0, v);
#pragma unroll
                for (int pt = 0; pt < PB; ++pt) {
                    int l[L];
                    split_limbs<L, W>(v[pt], l);
#pragma unroll
                    for (int t = 0; t < L; ++t)
                        words[pt][t] |= (uint32_t)(l[t] & 255) << (8 * u);
                }
            }
        }
#pragma unroll
        for (int pt = 0; pt < PB; ++pt)
#pragma unroll
            for (int t = 0; t < L; ++t)
                *reinterpret_cast<uint32_t *>(lb + ((size_t)(pt * L + t) * Np + j) * Kp + 4 * kq) =
                    words[pt][t];
    }
    __syncthreads();

    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int tilesM = Mp >> 5, tilesN = Np >> 5;
    const int ntask = PB * tilesM * tilesN;
    constexpr int NG = (L == 8) ? 3 : 2;   // partial sums
    constexpr int GW = (L == 8) ? 5 : 4;   // weights per partial sum (last one may be shorter)

    for (int task = wave; task < ntask; task += kThreads / 64) {
        const int pt = task / (tilesM * tilesN);
        const int rem = task - pt * tilesM * tilesN;
        const int mt = rem / tilesN;
        const int nt = rem - mt * tilesN;

        int64_t G[NG][16];
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
            for (int r = 0; r < 16; ++r)
                G[g][r] = 0;

        for (int kc = 0; kc < Kp; kc += 32) {
            v4i afr[L], bfr[L];
            const int koff = kc + 16 * (lane >> 5);
#pragma unroll
            for (int t = 0; t < L; ++t) {
                afr[t] = *reinterpret_cast<const v4i *>(
                    la + ((size_t)(pt * L + t) * Mp + mt * 32 + (lane & 31)) * Kp + koff);
                bfr[t] = *reinterpret_cast<const v4i *>(
                    lb + ((size_t)(pt * L + t) * Np + nt * 32 + (lane & 31)) * Kp + koff);
            }
#pragma unroll
            for (int w = 0; w < 2 * L - 1; ++w) {
                v16i acc = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int s = 0; s < L; ++s) {
                    const int t = w - s;
                    if (t >= 0 && t < L)
                        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(afr[s], bfr[t], acc, 0, 0, 0);
                }
                const int g = w / GW;
                const int sh = 8 * (w - g * GW);
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    G[g][r] += (int64_t)acc[r] << sh;
            }
        }

        const int col = nt * 32 + (lane & 31);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = mt * 32 + mfma_row(r, lane);
            uint64_t res;
            if constexpr (L == 8) {
                const uint64_t u0 = (uint64_t)(G[0][r] + (int64_t)mc.goff);
                const uint64_t u1 = (uint64_t)(G[1][r] + (int64_t)mc.goff);
                const uint64_t u2 = (uint64_t)(G[2][r] + (int64_t)mc.goff);
                uint64_t s = barrett_lazy(u0, mc.bm, mc.p) + shoup_lazy(u1, mc.c40, mc.c40q, mc.p) +
                             shoup_lazy(u2, mc.c80, mc.c80q, mc.p);
                res = barrett(s, mc.bm, mc.p);
            } else {
                const uint64_t u0 = (uint64_t)(G[0][r] + (int64_t)mc.goff);
                const uint64_t u1 = (uint64_t)(G[1][r] + (int64_t)mc.goff);
                uint64_t s = barrett_lazy(u0, mc.bm, mc.p) + shoup_lazy(u1, mc.c32, mc.c32q, mc.p);
                res = barrett(s, mc.bm, mc.p);
            }
            if (row < m && col < n)
                Ch[((int64_t)row * n + col) * N + q0 + pt] = (W)res;
        }
    }
}

inline int64_t mfma_point_bytes(int m, int k, int n, int limbs)
{
    const int64_t Mp = (m + 31) & ~31, Kp = (k + 31) & ~31, Np = (n + 31) & ~31;
    return (int64_t)limbs * (Mp * Kp + Np * Kp);
}

template <typename W, int L, int PB>
hipError_t launch_mfma_pb(const W *Ah, const W *Bh, W *Ch, int m, int k, int n, int64_t N,
                          const ModConst &mc, hipStream_t s)
{
    const size_t lds = (size_t)PB * mfma_point_bytes(m, k, n, L);
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)pw_mfma_kernel<W, L, PB>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBudget);
        if (e != hipSuccess)
            return e;
        attr_set = true;
    }
    const int64_t blocks = N / PB;
    hipLaunchKernelGGL((pw_mfma_kernel<W, L, PB>), dim3((unsigned)blocks), dim3(kThreads), lds, s,
                       Ah, Bh, Ch, m, k, n, N, mc);
    return hipGetLastError();
}

template <typename W, int L>
hipError_t launch_mfma(const W *Ah, const W *Bh, W *Ch, int m, int k, int n, int64_t N,
                       const ModConst &mc, hipStream_t s)
{
    const int64_t per = mfma_point_bytes(m, k, n, L);
    if (N >= 8 && 8 * per <= kLdsBudget)
        return launch_mfma_pb<W, L, 8>(Ah, Bh, Ch, m, k, n, N, mc, s);
    if (N >= 4 && 4 * per <= kLdsBudget)
        return launch_mfma_pb<W, L, 4>(Ah, Bh, Ch, m, k, n, N, mc, s);
    if (N >= 2 && 2 * per <= kLdsBudget)
        return launch_mfma_pb<W, L, 2>(Ah, Bh, Ch, m, k, n, N, mc, s);
    return launch_mfma_pb<W, L, 1>(Ah, Bh, Ch, m, k, n, N, mc, s);
}

}  // namespace

bool mfma_supported(int m, int k, int n, int limbs)
{
    // per-point limb planes must fit the LDS budget; the signed partial sums
    // stay below 2^61 for K <= 4064 (|G| < 2^54.01 per 32-deep chunk)
    return m > 0 && k > 0 && n > 0 && k <= 4064 && mfma_point_bytes(m, k, n, limbs) <= kLdsBudget;
}

template <>
hipError_t launch_pointwise<uint64_t>(PointwiseKind kind, const uint64_t *Ah, const uint64_t *Bh,
                                      uint64_t *Ch, int m, int k, int n, int64_t N,
                                      const ModConst &mc, hipStream_t s)
{
    if (kind == PW_MFMA) {
        if (mc.limbs != 8 || !mfma_supported(m, k, n, 8))
            return hipErrorInvalidValue;
        return launch_mfma<uint64_t, 8>(Ah, Bh, Ch, m, k, n, N, mc, s);
    }
    const int tiles = ((m + TM - 1) / TM) * ((n + TN - 1) / TN);
    dim3 grid((unsigned)((N + 63) / 64), (unsigned)((tiles + 3) / 4));
    hipLaunchKernelGGL(pw_plain_u64, grid, dim3(kThreads), 0, s, Ah, Bh, Ch, m, k, n, N, mc);
    return hipGetLastError();
}

template <>
hipError_t launch_pointwise<uint32_t>(PointwiseKind kind, const uint32_t *Ah, const uint32_t *Bh,
                                      uint32_t *Ch, int m, int k, int n, int64_t N,
                                      const ModConst &mc, hipStream_t s)
{
    if (kind == PW_MFMA) {
        if (mc.limbs != 4 || !mfma_supported(m, k, n, 4))
            return hipErrorInvalidValue;
        return launch_mfma<uint32_t, 4>(Ah, Bh, Ch, m, k, n, N, mc, s);
    }
    const int tiles = ((m + TM - 1) / TM) * ((n + TN - 1) / TN);
    dim3 grid((unsigned)((N + 63) / 64), (unsigned)((tiles + 3) / 4));
    hipLaunchKernelGGL(pw_plain_u32, grid, dim3(kThreads), 0, s, Ah, Bh, Ch, m, k, n, N, mc);
    return hipGetLastError();
}

}  // namespace pmlg
