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
//    three (or two) signed 64-bit partial sums by shifts, the exact dot
//    product is rebuilt in 128 bits and reduced: by Montgomery REDC for 8
//    limbs (output V 2^-64 mod p in [0, 4p), the inverse transform's scaling
//    carries the 2^64), by Barrett to [0, p) for 4 limbs.  No floating point
//    is involved.
#include "internal.hpp"

namespace pmlg {

namespace {

constexpr int kThreads = 256;
constexpr int kLdsBudget = 128 * 1024;  // bytes of LDS for limb planes

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// Point-blocked evaluation layout (see launch_ntt_fwd): entry e of `count`
// at point q is element ((q >> logP) * count + e) * 2^logP + (q mod 2^logP).
__device__ __forceinline__ int64_t blk_addr(int64_t e, int64_t q, int64_t count, int logP)
{
    return (((q >> logP) * count + e) << logP) | (q & ((1 << logP) - 1));
}

// ---------------------------------------------------------------------------
// plain kernel
// ---------------------------------------------------------------------------
constexpr int TM = 2, TN = 4;

__device__ __forceinline__ void plain_fold(uint64_t &hi, uint64_t &lo, const ModConst &mc)
{
    lo = reduce128(hi, lo, mc);
    hi = 0;
}

// Lane (point q, output tile `tile` of TM x TN entries) of the plain product.
__device__ __forceinline__ void plain_u64_body(const uint64_t *__restrict__ Ah,
                                               const uint64_t *__restrict__ Bh,
                                               uint64_t *__restrict__ Ch, int m, int k, int n,
                                               int64_t N, int logP, int logPc, const ModConst &mc,
                                               int64_t q, int tile)
{
    const int tiles_n = (n + TN - 1) / TN;
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
            a[r] = (i0 + r < m) ? Ah[blk_addr((int64_t)(i0 + r) * k + kk, q, (int64_t)m * k, logP)] : 0;
#pragma unroll
        for (int c = 0; c < TN; ++c)
            b[c] = (j0 + c < n) ? Bh[blk_addr((int64_t)kk * n + j0 + c, q, (int64_t)k * n, logP)] : 0;
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
                Ch[blk_addr((int64_t)(i0 + r) * n + j0 + c, q, (int64_t)m * n, logPc)] =
                    reduce128(hi[r][c], lo[r][c], mc);
}

__device__ __forceinline__ void plain_u32_body(const uint32_t *__restrict__ Ah,
                                               const uint32_t *__restrict__ Bh,
                                               uint32_t *__restrict__ Ch, int m, int k, int n,
                                               int64_t N, int logP, int logPc, const ModConst &mc,
                                               int64_t q, int tile)
{
    const int tiles_n = (n + TN - 1) / TN;
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
            a[r] = (i0 + r < m) ? Ah[blk_addr((int64_t)(i0 + r) * k + kk, q, (int64_t)m * k, logP)] : 0;
#pragma unroll
        for (int c = 0; c < TN; ++c)
            b[c] = (j0 + c < n) ? Bh[blk_addr((int64_t)kk * n + j0 + c, q, (int64_t)k * n, logP)] : 0;
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
                Ch[blk_addr((int64_t)(i0 + r) * n + j0 + c, q, (int64_t)m * n, logPc)] =
                    (uint32_t)barrett(acc[r][c], mc.bm, mc.p);
}

// grid: x = 64-point chunks, y = groups of 4 output tiles (one per wave)
template <typename W>
__global__ void __launch_bounds__(kThreads)
pw_plain(const W *__restrict__ Ah, const W *__restrict__ Bh, W *__restrict__ Ch, int m, int k,
         int n, int64_t N, int logP, int logPc, ModConst mc)
{
    const int64_t q = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
    const int tile = blockIdx.y * 4 + (threadIdx.x >> 6);
    if constexpr (sizeof(W) == 8)
        plain_u64_body(Ah, Bh, Ch, m, k, n, N, logP, logPc, mc, q, tile);
    else
        plain_u32_body(Ah, Bh, Ch, m, k, n, N, logP, logPc, mc, q, tile);
}

// ---------------------------------------------------------------------------
// int8-limb MFMA kernel
// ---------------------------------------------------------------------------

// Balanced signed base-256 digits: with C = 0x80..80 (one 0x80 per byte),
// y = x + C has ordinary digits b_t and x = sum (b_t - 128) 256^t, so the
// signed digit l_t = b_t - 128 is the byte b_t ^ 0x80, i.e. byte t of
// (x + C) ^ C.  Valid while x + C does not overflow (x < 2^61 for 8 limbs,
// x < 2^30 for 4 limbs).
__device__ __forceinline__ uint64_t limb_bytes64(uint64_t x)
{
    const uint64_t C = 0x8080808080808080ull;
    return (x + C) ^ C;
}
__device__ __forceinline__ uint32_t limb_bytes32(uint32_t x)
{
    const uint32_t C = 0x80808080u;
    return (x + C) ^ C;
}

// 4x4 byte transpose: out[t] = byte t of in[0] | byte t of in[1] << 8 | ...
__device__ __forceinline__ void byte_transpose4(const uint32_t (&in)[4], uint32_t (&out)[4])
{
    // v_perm_b32(s0, s1, sel): selector 0-3 picks bytes of s1, 4-7 bytes of s0
    const uint32_t t0 = __builtin_amdgcn_perm(in[1], in[0], 0x06020400u);
    const uint32_t t1 = __builtin_amdgcn_perm(in[1], in[0], 0x07030501u);
    const uint32_t t2 = __builtin_amdgcn_perm(in[3], in[2], 0x06020400u);
    const uint32_t t3 = __builtin_amdgcn_perm(in[3], in[2], 0x07030501u);
    out[0] = __builtin_amdgcn_perm(t2, t0, 0x05040100u);
    out[2] = __builtin_amdgcn_perm(t2, t0, 0x07060302u);
    out[1] = __builtin_amdgcn_perm(t3, t1, 0x05040100u);
    out[3] = __builtin_amdgcn_perm(t3, t1, 0x07060302u);
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

// Loads 4 entries (given by their offsets) x PB points and writes the L limb
// planes of each point as 32-bit words (4 consecutive k) at dst + (pt*L+t)*plane.
template <typename W, int L, int PB>
__device__ __forceinline__ void pack_quad(const W *__restrict__ base, const int64_t (&off)[4],
                                          const bool (&ok)[4], unsigned char *dst, size_t plane)
{
    W v[4][PB];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (ok[u]) {
            load_run<W, PB>(base + off[u], v[u]);
        } else {
#pragma unroll
            for (int pt = 0; pt < PB; ++pt)
                v[u][pt] = 0;
        }
    }
#pragma unroll
    for (int pt = 0; pt < PB; ++pt) {
        if constexpr (L == 8) {
            uint32_t lo[4], hi[4], wlo[4], whi[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const uint64_t z = limb_bytes64((uint64_t)v[u][pt]);
                lo[u] = (uint32_t)z;
                hi[u] = (uint32_t)(z >> 32);
            }
            byte_transpose4(lo, wlo);
            byte_transpose4(hi, whi);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                *reinterpret_cast<uint32_t *>(dst + (size_t)(pt * L + t) * plane) = wlo[t];
                *reinterpret_cast<uint32_t *>(dst + (size_t)(pt * L + 4 + t) * plane) = whi[t];
            }
        } else {
            uint32_t in[4], w[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                in[u] = limb_bytes32((uint32_t)v[u][pt]);
            byte_transpose4(in, w);
#pragma unroll
            for (int t = 0; t < 4; ++t)
                *reinterpret_cast<uint32_t *>(dst + (size_t)(pt * L + t) * plane) = w[t];
        }
    }
}

// Output layout of v_mfma_i32_32x32x32_i8: register r of lane l holds
// D[row][col] with col = l % 32 and row = 8 (r / 4) + 4 (l / 32) + r % 4.
__device__ __forceinline__ int mfma_row(int r, int lane)
{
    return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}

// Point block `blk` (points blk*PB .. blk*PB+PB-1) of the MFMA product, one
// workgroup; the caller separates two calls on one workgroup by a barrier.
template <typename W, int L, int PB>
__device__ __forceinline__ void mfma_body(const W *__restrict__ Ah, const W *__restrict__ Bh,
                                          W *__restrict__ Ch, int m, int k, int n, int logP,
                                          int logPc, const ModConst &mc, int64_t blk)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
    const int Mp = (m + 31) & ~31, Kp = (k + 31) & ~31, Np = (n + 31) & ~31;
    const int Kq = Kp >> 2;
    unsigned char *la = lds;                             // [PB][L][Mp][Kp]
    unsigned char *lb = lds + (size_t)PB * L * Mp * Kp;  // [PB][L][Np][Kp]  (B transposed)
    const int64_t q0 = blk * PB;
    const int tid = threadIdx.x;

    // ---- A limbs: one thread per (row i, 4 consecutive k) ----
    for (int quad = tid; quad < Mp * Kq; quad += kThreads) {
        const int i = quad / Kq;
        const int kq = quad - i * Kq;
        int64_t off[4];
        bool ok[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kk = 4 * kq + u;
            ok[u] = i < m && kk < k;
            off[u] = blk_addr((int64_t)i * k + kk, q0, (int64_t)m * k, logP);
        }
        pack_quad<W, L, PB>(Ah, off, ok, la + (size_t)i * Kp + 4 * kq, (size_t)Mp * Kp);
    }
    // ---- B limbs, transposed: one thread per (column j, 4 consecutive k) ----
    for (int quad = tid; quad < Np * Kq; quad += kThreads) {
        const int j = quad / Kq;
        const int kq = quad - j * Kq;
        int64_t off[4];
        bool ok[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int kk = 4 * kq + u;
            ok[u] = j < n && kk < k;
            off[u] = blk_addr((int64_t)kk * n + j, q0, (int64_t)k * n, logP);
        }
        pack_quad<W, L, PB>(Bh, off, ok, lb + (size_t)j * Kp + 4 * kq, (size_t)Np * Kp);
    }
    __syncthreads();

    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int tilesM = Mp >> 5, tilesN = Np >> 5;
    const int ntask = PB * tilesM * tilesN;
    constexpr int NG = (L == 8) ? 3 : 2;   // partial sums
    constexpr int GW = (L == 8) ? 5 : 4;   // weights per partial sum (last one may be shorter)

    // output addressing: entry (row, col) of point q0+pt is at
    // cbase + (row*n + col) * 2^logPc with cbase fixed by the point block
    // (PB <= 2^logP <= 2^logPc: the block lies within one run of C)
    const int64_t cbase = (((q0 >> logPc) * ((int64_t)m * n)) << logPc) | (q0 & ((1 << logPc) - 1));

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
            // A fragments stay in registers; B fragments are re-read from LDS
            // at each use (keeps the wave within 256 registers)
            v4i afr[L];
            const int koff = kc + 16 * (lane >> 5);
            const unsigned char *bbase = lb + ((size_t)(pt * L) * Np + nt * 32 + (lane & 31)) * Kp + koff;
            const size_t bplane = (size_t)Np * Kp;
#pragma unroll
            for (int t = 0; t < L; ++t)
                afr[t] = *reinterpret_cast<const v4i *>(
                    la + ((size_t)(pt * L + t) * Mp + mt * 32 + (lane & 31)) * Kp + koff);
#define BFR(t) (*reinterpret_cast<const v4i *>(bbase + (size_t)(t) * bplane))
            // weights w and w+1 of one partial sum are combined in 32 bits first:
            // |T_w + 256 T_(w+1)| <= 2^22 + 2^30 < 2^31
#pragma unroll
            for (int w = 0; w < 2 * L - 1; ++w) {
                const int g = w / GW;
                const int sh = w - g * GW;
                const bool pair = (sh % 2 == 0) && (w + 1 < 2 * L - 1) && ((w + 1) / GW == g);
                if (sh % 2 == 1)
                    continue;  // consumed as the upper half of a pair
                v16i acc0 = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int s2 = 0; s2 < L; ++s2) {
                    const int t2 = w - s2;
                    if (t2 >= 0 && t2 < L)
                        acc0 = __builtin_amdgcn_mfma_i32_32x32x32_i8(afr[s2], BFR(t2), acc0, 0, 0, 0);
                }
                if (pair) {
                    v16i acc1 = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
                    for (int s2 = 0; s2 < L; ++s2) {
                        const int t2 = w + 1 - s2;
                        if (t2 >= 0 && t2 < L)
                            acc1 = __builtin_amdgcn_mfma_i32_32x32x32_i8(afr[s2], BFR(t2), acc1, 0, 0, 0);
                    }
#pragma unroll
                    for (int r = 0; r < 16; ++r)
                        acc0[r] = acc0[r] + (int)((uint32_t)acc1[r] << 8);
                }
#pragma unroll
                for (int r = 0; r < 16; ++r)
                    G[g][r] += (int64_t)acc0[r] * ((int64_t)1 << (8 * sh));
            }
#undef BFR
        }

        const int col = nt * 32 + (lane & 31);
        const int row0 = mt * 32 + 4 * (lane >> 5);
        const int nP = n << logPc;  // row stride of C in the blocked layout (uniform, 32-bit)
        W *cp = Ch + cbase + pt + (((int64_t)row0 * n + col) << logPc);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int dr = 8 * (r >> 2) + (r & 3);  // mfma_row(r, lane) - 4 (lane / 32)
            W res;
            if constexpr (L == 8) {
                // exact V = G0 + G1 2^40 + G2 2^80 as a 128-bit integer, 0 <= V < 2^128,
                // then Montgomery REDC: t = (V + m p) / 2^64 = V 2^-64 mod p, t < hi + p + 1
                // (the 2^-64 is undone by the inverse transform's scaling)
                const int64_t g0 = G[0][r], g1 = G[1][r], g2 = G[2][r];
                const uint64_t lo = (uint64_t)g0 + ((uint64_t)g1 << 40);
                const uint64_t carry = lo < (uint64_t)g0 ? 1u : 0u;
                const uint64_t hi = (uint64_t)(g0 >> 63) + (uint64_t)(g1 >> 24) +
                                    ((uint64_t)g2 << 16) + carry;
                const uint64_t mm = lo * mc.pinvneg;
                uint64_t t = hi + mulhi64(mm, mc.p) + (lo != 0 ? 1u : 0u);
                // small k p: t < 4p already; otherwise a Barrett step (t < 2^64)
                if (!mc.redc_small)
                    t = barrett_lazy(t, mc.bm, mc.p);
                res = (W)t;
            } else {
                // V = G0 + G1 2^32 < k p^2 < 2^72: hi < 2^8, hi * (2^64 mod p) < 2^38
                const int64_t g0 = G[0][r], g1 = G[1][r];
                const uint64_t lo = (uint64_t)g0 + ((uint64_t)g1 << 32);
                const uint64_t carry = lo < (uint64_t)g0 ? 1u : 0u;
                const uint64_t hi = (uint64_t)(g0 >> 63) + (uint64_t)(g1 >> 32) + carry;
                res = (W)barrett(barrett_lazy(lo, mc.bm, mc.p) + hi * mc.c64, mc.bm, mc.p);
            }
            if (row0 + dr < m && col < n)
                cp[dr * nP] = res;
        }
    }
}

template <typename W, int L, int PB>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
pw_mfma_kernel(const W *__restrict__ Ah, const W *__restrict__ Bh, W *__restrict__ Ch, int m,
               int k, int n, int64_t N, int logP, int logPc, ModConst mc)
{
    // XCD-aware order: workgroups b, b+8, b+16, ... run on one XCD (round-robin
    // dispatch), so give them consecutive point blocks to share its L2
    const int64_t nblk = N / PB;
    int64_t blk = blockIdx.x;
    if ((nblk & 7) == 0)
        blk = (blk & 7) * (nblk >> 3) + (blk >> 3);
    mfma_body<W, L, PB>(Ah, Bh, Ch, m, k, n, logP, logPc, mc, blk);
}

// ---------------------------------------------------------------------------
// persistent segment-table launches (batched products)
// ---------------------------------------------------------------------------
template <typename W>
__device__ __forceinline__ const PwSeg<W> &pw_seg_of(const PwSeg<W> *segs, int nseg, int64_t b)
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

// Every workgroup walks the task list b = blockIdx.x, + gridDim.x, ... until
// `blocks`; a task is one point block of one product.
template <typename W, int L, int PB>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2, 2)))
pw_mfma_segs(const PwSeg<W> *__restrict__ segs, int nseg, int64_t blocks, ModConst mc)
{
    for (int64_t b = blockIdx.x; b < blocks; b += gridDim.x) {
        const PwSeg<W> &sg = pw_seg_of(segs, nseg, b);
        ModConst mcs = mc;
        mcs.redc_small = sg.redc_small;
        mfma_body<W, L, PB>(sg.Ah, sg.Bh, sg.Ch, sg.m, sg.k, sg.n, sg.logP, sg.logPc, mcs,
                            b - sg.block0);
        __syncthreads();  // the next task rewrites the limb planes
    }
}

// task = (64-point chunk, group of 4 output tiles) of one product
template <typename W>
__global__ void __launch_bounds__(kThreads)
pw_plain_segs(const PwSeg<W> *__restrict__ segs, int nseg, int64_t blocks, ModConst mc)
{
    for (int64_t b = blockIdx.x; b < blocks; b += gridDim.x) {
        const PwSeg<W> &sg = pw_seg_of(segs, nseg, b);
        const int64_t gx = (sg.N + 63) >> 6;
        const int64_t lb = b - sg.block0;
        const int64_t by = lb / gx;
        const int64_t q = (lb - by * gx) * 64 + (threadIdx.x & 63);
        const int tile = (int)by * 4 + (threadIdx.x >> 6);
        if constexpr (sizeof(W) == 8)
            plain_u64_body(sg.Ah, sg.Bh, sg.Ch, sg.m, sg.k, sg.n, sg.N, sg.logP, sg.logPc, mc, q, tile);
        else
            plain_u32_body(sg.Ah, sg.Bh, sg.Ch, sg.m, sg.k, sg.n, sg.N, sg.logP, sg.logPc, mc, q, tile);
    }
}

inline int64_t mfma_point_bytes(int m, int k, int n, int limbs)
{
    const int64_t Mp = (m + 31) & ~31, Kp = (k + 31) & ~31, Np = (n + 31) & ~31;
    return (int64_t)limbs * (Mp * Kp + Np * Kp);
}

template <typename W, int L, int PB>
hipError_t launch_mfma_pb(const W *Ah, const W *Bh, W *Ch, int m, int k, int n, int64_t N,
                          int logP, int logPc, const ModConst &mc, hipStream_t s)
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
                       Ah, Bh, Ch, m, k, n, N, logP, logPc, mc);
    return hipGetLastError();
}

template <typename W, int L>
hipError_t launch_mfma(const W *Ah, const W *Bh, W *Ch, int m, int k, int n, int64_t N, int logP,
                       int logPc, const ModConst &mc, hipStream_t s)
{
    // Points per workgroup: within one layout block, and at most 64 KiB of
    // limb planes so that two workgroups share a CU.
    const int64_t per = mfma_point_bytes(m, k, n, L);
    constexpr int64_t pref = 64 * 1024;
    const int64_t P = (int64_t)1 << logP;
    if (P >= 8 && 8 * per <= pref)
        return launch_mfma_pb<W, L, 8>(Ah, Bh, Ch, m, k, n, N, logP, logPc, mc, s);
    if (P >= 4 && 4 * per <= pref)
        return launch_mfma_pb<W, L, 4>(Ah, Bh, Ch, m, k, n, N, logP, logPc, mc, s);
    if (P >= 2 && 2 * per <= pref)
        return launch_mfma_pb<W, L, 2>(Ah, Bh, Ch, m, k, n, N, logP, logPc, mc, s);
    return launch_mfma_pb<W, L, 1>(Ah, Bh, Ch, m, k, n, N, logP, logPc, mc, s);
}

}  // namespace

bool mfma_supported(int m, int k, int n, int limbs, uint64_t p)
{
    // per-point limb planes must fit the LDS budget; the signed partial sums
    // stay below 2^61 for K <= 4064 (|G| < 2^54.01 per 32-deep chunk); the
    // exact recombination needs the dot product k (p-1)^2 below 2^128
    if (m <= 0 || k <= 0 || n <= 0 || k > 4064 || mfma_point_bytes(m, k, n, limbs) > kLdsBudget)
        return false;
    // REDC needs V + (p+1) 2^64 <= 2^128 (8 limbs); 4 limbs only needs V < 2^72
    const unsigned __int128 sq = (unsigned __int128)(p - 1) * (p - 1);
    const unsigned __int128 lim =
        limbs == 8 ? (~(unsigned __int128)0) - ((unsigned __int128)(p + 1) << 64)
                   : ((unsigned __int128)1 << 72);
    return sq == 0 || (unsigned __int128)k <= lim / sq;
}

template <>
hipError_t launch_pointwise<uint64_t>(PointwiseKind kind, const uint64_t *Ah, const uint64_t *Bh,
                                      uint64_t *Ch, int m, int k, int n, int64_t N, int logP,
                                      int logPc, const ModConst &mc, hipStream_t s)
{
    if (kind == PW_MFMA) {
        if (mc.limbs != 8 || !mfma_supported(m, k, n, 8, mc.p))
            return hipErrorInvalidValue;
        return launch_mfma<uint64_t, 8>(Ah, Bh, Ch, m, k, n, N, logP, logPc, mc, s);
    }
    const int tiles = ((m + TM - 1) / TM) * ((n + TN - 1) / TN);
    dim3 grid((unsigned)((N + 63) / 64), (unsigned)((tiles + 3) / 4));
    hipLaunchKernelGGL(pw_plain<uint64_t>, grid, dim3(kThreads), 0, s, Ah, Bh, Ch, m, k, n, N,
                       logP, logPc, mc);
    return hipGetLastError();
}

template <>
hipError_t launch_pointwise<uint32_t>(PointwiseKind kind, const uint32_t *Ah, const uint32_t *Bh,
                                      uint32_t *Ch, int m, int k, int n, int64_t N, int logP,
                                      int logPc, const ModConst &mc, hipStream_t s)
{
    if (kind == PW_MFMA) {
        if (mc.limbs != 4 || !mfma_supported(m, k, n, 4, mc.p))
            return hipErrorInvalidValue;
        return launch_mfma<uint32_t, 4>(Ah, Bh, Ch, m, k, n, N, logP, logPc, mc, s);
    }
    const int tiles = ((m + TM - 1) / TM) * ((n + TN - 1) / TN);
    dim3 grid((unsigned)((N + 63) / 64), (unsigned)((tiles + 3) / 4));
    hipLaunchKernelGGL(pw_plain<uint32_t>, grid, dim3(kThreads), 0, s, Ah, Bh, Ch, m, k, n, N,
                       logP, logPc, mc);
    return hipGetLastError();
}


// ---------------------------------------------------------------------------
// batched launches
// ---------------------------------------------------------------------------
int64_t pw_seg_blocks(PointwiseKind kind, int pb, int m, int n, int64_t N)
{
    if (kind == PW_MFMA)
        return N / pb;
    const int tiles = ((m + TM - 1) / TM) * ((n + TN - 1) / TN);
    return ((N + 63) / 64) * ((tiles + 3) / 4);
}

size_t pw_mfma_lds(int pb, int m, int k, int n, int limbs)
{
    return (size_t)pb * mfma_point_bytes(m, k, n, limbs);
}

int pw_pick_pb(int m, int k, int n, int limbs, int logP)
{
    const int64_t per = mfma_point_bytes(m, k, n, limbs);
    constexpr int64_t pref = 64 * 1024;
    for (int pb = 8; pb > 1; pb >>= 1)
        if (pb <= (1 << logP) && pb * per <= pref)
            return pb;
    return 1;
}

namespace {
template <typename W, int L, int PB>
hipError_t mfma_segs_pb(const PwSeg<W> *segs, int nseg, int64_t blocks, size_t lds, int grid,
                        const ModConst &mc, hipStream_t s)
{
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void *)pw_mfma_segs<W, L, PB>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBudget);
        if (e != hipSuccess)
            return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((pw_mfma_segs<W, L, PB>), dim3((unsigned)grid), dim3(kThreads), lds, s, segs,
                       nseg, blocks, mc);
    return hipGetLastError();
}
}  // namespace

template <typename W>
hipError_t launch_pointwise_segs(PointwiseKind kind, int pb, const PwSeg<W> *segs, int nseg,
                                 int64_t blocks, size_t lds, int grid, const ModConst &mc,
                                 hipStream_t s)
{
    if (nseg <= 0 || blocks <= 0)
        return hipSuccess;
    if (grid <= 0 || grid > blocks)
        grid = (int)(blocks < (1 << 30) ? blocks : (1 << 30));
    if (kind == PW_PLAIN) {
        hipLaunchKernelGGL(pw_plain_segs<W>, dim3((unsigned)grid), dim3(kThreads), 0, s, segs, nseg,
                           blocks, mc);
        return hipGetLastError();
    }
    constexpr int L = sizeof(W) == 8 ? 8 : 4;
    if (mc.limbs != L || lds > (size_t)kLdsBudget)
        return hipErrorInvalidValue;
    switch (pb) {
    case 8:
        return mfma_segs_pb<W, L, 8>(segs, nseg, blocks, lds, grid, mc, s);
    case 4:
        return mfma_segs_pb<W, L, 4>(segs, nseg, blocks, lds, grid, mc, s);
    case 2:
        return mfma_segs_pb<W, L, 2>(segs, nseg, blocks, lds, grid, mc, s);
    case 1:
        return mfma_segs_pb<W, L, 1>(segs, nseg, blocks, lds, grid, mc, s);
    default:
        return hipErrorInvalidValue;
    }
}

template hipError_t launch_pointwise_segs<uint64_t>(PointwiseKind, int, const PwSeg<uint64_t> *,
                                                    int, int64_t, size_t, int, const ModConst &,
                                                    hipStream_t);
template hipError_t launch_pointwise_segs<uint32_t>(PointwiseKind, int, const PwSeg<uint32_t> *,
                                                    int, int64_t, size_t, int, const ModConst &,
                                                    hipStream_t);

}  // namespace pmlg
