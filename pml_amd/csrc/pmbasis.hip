// Approximant bases on the device: M-Basis (iterative, one order per step)
// and PM-Basis (divide and conquer on the order, with the device middle
// product and product), following
//   flint-extras/src/nmod_mat_poly_extra/nmod_mat_poly_mbasis.c:116-378
//   flint-extras/src/nmod_poly_mat_extra/approximant_basis.c:18-69
//   flint-extras/src/nmod_mat_extra/left_nullspace.c:42-102
//
// One M-Basis step at order `ord`:
//  1. residual = coefficient ord of appbas * F.  It is kept as a truncated
//     series R = appbas * F mod x^order, updated by the same row operations
//     as appbas (the value is unique, so this equals the reference's
//     nmod_mat_poly_mul_coeff);
//  2. perm = stable sort of (shift[i], i); residual rows permuted by it;
//  3. left nullspace of the permuted residual in compact form: from the
//     reduced row echelon form R of its transpose (computed fraction-free,
//     then each pivot row divided by its pivot), X[t][r] = -R[r][nonpivot_t],
//     permutation = (pivot columns) ++ (non-pivot columns);
//  4. nullity 0: appbas <- x^(order-ord) appbas, shift += order-ord, stop;
//     nullity m: nothing; otherwise with piv[i] = perm[permutation[i]]:
//     shift[piv[r]] += 1 (r < rank), row piv[rank+t] += sum_r X[t][r] row piv[r],
//     rows piv[r] multiplied by x (length grows when one of them reaches the top).
//
// Kernel 1 (one workgroup) does 2-4 and writes the step's description;
// kernel 2 (a grid) applies the row operations to appbas and R, reading one
// buffer and writing the other.  All control (nullity cases, basis length)
// stays on the device: a leaf is a fixed sequence of launches.  The steps
// skip the division by the pivots in 3: with D the product of the pivots,
// they apply D X and multiply the updated rows by D, so each basis row is the
// reference's row times a known nonzero scale, and the leaf divides the
// scales out at its end.  The step's decisions only read which entries are
// zero, which row scaling keeps, and the normalised kernel of the scaled
// residual, applied to the scaled rows, gives the scaled normalised update.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <type_traits>

#include <climits>

#include "ctx.hpp"
#include "host_arith.hpp"

namespace pmlg {

namespace {

constexpr int kStepThreads = 1024;
constexpr int kMaxM = 256;    // rows of F (basis dimension) in the device M-Basis step
constexpr int kMaxN = 128;    // columns of F: at most two rows of the transposed residual per lane
constexpr int kPanel = 4;     // widest Gauss-Jordan panel (LDS is sized for it)
// dynamic LDS of the step kernel, set once for every instance: the value the
// 128 x 128 shape needs, which every allowed shape must fit
constexpr size_t kStepLdsMax = 147968;

// LDS bytes of the step kernel for an m x n residual with words of wbytes:
// shifts (m 64-bit words), At (n rows of odd stride m | 1), the pivot
// cofactors (kMaxN), two coefficient buffers (kMaxN x (kPanel + 1) each),
// then the int arrays (4 of m, 3 of kMaxN); as laid out in the kernel
__host__ __device__ inline size_t step_lds_bytes(int m, int n, size_t wbytes)
{
    return (size_t)m * 8 +
           ((size_t)n * (m | 1) + kMaxN + 2 * (size_t)kMaxN * (kPanel + 1)) * wbytes +
           (4 * (size_t)m + 3 * (size_t)kMaxN) * sizeof(int);
}
constexpr int kXPad = 32;     // words after X read (and discarded) by the last update group

// control words of one leaf
enum Ctl {
    C_CUR = 0,    // index (0/1) of the buffers holding the current appbas and R
    C_MODE = 1,   // what kernel 2 does at this step: 0 nothing, 1 update, 2 shift all
    C_RANK = 2,   // rank of the residual at this step
    C_LEN = 3,    // current length of appbas
    C_GROW = 4,   // set by kernel 2 when the length grows by one
    C_DONE = 5,   // full-rank residual met: the leaf is finished
    C_SHIFT = 6,  // mode 2: power of x
    C_LOLD = 7,   // mode 2: length before the shift
    C_FLIP = 8,   // kernel 2 wrote the other buffers: swap before the next step
    C_WORDS = 16
};

struct LeafArgs {
    int m, n;
    int ordL;     // order of the leaf
    int cap;      // coefficient capacity of appbas (ordL + 1)
    uint64_t *app[2];  // m x m x cap, entry-major
    uint64_t *R[2];    // m x n x ordL, entry-major
    uint64_t *Rn;      // m x n: the coming step's residual coefficient
    // Row scales.  The steps use unnormalised kernel vectors (own coefficient
    // D, the product of the step's pivots, instead of 1), so basis row i is
    // scale[i] times the normalised one; the leaf divides them out at its
    // end.  scale[m] holds the current step's D for the update (plain residue
    // for 64-bit words, Montgomery residue D 2^32 mod p for 32-bit ones).
    uint64_t *scale;   // m + 1 words
    uint64_t *X;       // nullity x rank, row stride m
    int *piv;          // m
    int *rowinfo;      // m: t >= 0 nullity row t, -1-r rank row r
    int64_t *shift;    // m
    int *ctl;          // C_WORDS
    ModConst mc;
};

__device__ __forceinline__ uint64_t mulmod_g(uint64_t a, uint64_t b, const ModConst &mc)
{
    if (mc.p < ((uint64_t)1 << 32))
        return barrett(a * b, mc.bm, mc.p);
    if (mc.bits > 62)
        return mulmod_any(a, b, mc);
    uint64_t hi, lo;
    mul128(a, b, hi, lo);
    return reduce128(hi, lo, mc);
}

__device__ __forceinline__ uint64_t submod_g(uint64_t a, uint64_t b, uint64_t p)
{
    return a >= b ? a - b : a + (p - b);
}

__device__ uint64_t invmod_g(uint64_t a, const ModConst &mc)
{
    // Fermat: a^(p-2)
    uint64_t e = mc.p - 2, r = 1, b = a;
    while (e) {
        if (e & 1)
            r = mulmod_g(r, b, mc);
        b = mulmod_g(b, b, mc);
        e >>= 1;
    }
    return r;
}

// Inverse of a in [1, p) modulo a prime p < 2^62 by the extended Euclidean
// algorithm, r_i = s_i a (mod p), with the quotients estimated in double
// precision.  f = fl(r0) / fl(r1) has relative error below 2^-51, so for
// f < 2^49 trunc(f) is within one of the quotient and one correction step
// fixes it; larger quotients (r1 tiny) are refined once on the remainder,
// which is then below (2^11 + 1) r1 in magnitude.  The cofactors stay below p
// in magnitude.  About 0.84 ln p division steps on average against ~log2 p
// dependent squarings for Fermat's a^(p-2).
__device__ uint64_t inv_euclid(uint64_t a, uint64_t p)
{
    uint64_t r0 = p, r1 = a;
    int64_t s0 = 0, s1 = 1;
    double d0 = (double)r0, d1 = (double)r1;  // fl(r0), fl(r1), carried along
    while (r1 != 0) {
        const double f = d0 / d1;
        uint64_t q = (uint64_t)f;
        int64_t rem = (int64_t)(r0 - q * r1);
        if (f >= 0x1p49) {
            const int64_t q2 = (int64_t)((double)rem / (double)r1);
            q += (uint64_t)q2;
            rem -= q2 * (int64_t)r1;
        }
        if (rem < 0) {
            rem += (int64_t)r1;
            --q;
        } else if ((uint64_t)rem >= r1) {
            rem -= (int64_t)r1;
            ++q;
        }
        const int64_t s2 = s0 - (int64_t)q * s1;
        r0 = r1;
        d0 = d1;
        r1 = (uint64_t)rem;
        d1 = (double)r1;
        s0 = s1;
        s1 = s2;
    }
    return s0 < 0 ? (uint64_t)(s0 + (int64_t)p) : (uint64_t)s0;
}

// Arithmetic of the step kernel (p an odd prime below 2^62).  Values in LDS
// are Montgomery residues: x 2^32 mod p for p < 2^30 (Mont32: one 32x32 -> 64
// product, one low product and one multiply-add per reduction), x 2^64 mod p
// otherwise (Mont64: a 64x64 -> 128 product, one low and one high 64-bit
// product, against two Barrett steps and a Shoup step for a plain 128-bit
// reduction).  Zero maps to zero, so pivot tests read the representation
// directly.
struct Mont64Arith {
    using V = uint64_t;
    uint64_t p, pinv, r1, r2;  // pinv = -p^-1 mod 2^64, r1 = 2^64 mod p, r2 = 2^128 mod p
    __device__ explicit Mont64Arith(const ModConst &c)
        : p(c.p), pinv(c.pinvneg), r1(c.r64), r2(mulmod_g(c.r64, c.r64, c)) {}
    __device__ V redc(uint64_t hi, uint64_t lo) const
    {
        // (hi 2^64 + lo + mm p) / 2^64 with mm = lo pinv mod 2^64: the low
        // words sum to 0 when lo = 0 and to 2^64 otherwise; for inputs below
        // p^2 the quotient is below 2p < 2^63
        const uint64_t mm = lo * pinv;
        const uint64_t u = hi + mulhi64(mm, p) + (lo != 0 ? 1 : 0);
        return u >= p ? u - p : u;
    }
    __device__ V mul(V a, V b) const
    {
        uint64_t hi, lo;
        mul128(a, b, hi, lo);
        return redc(hi, lo);
    }
    // Sums of up to five products, reduced once: t < 5 p^2 < 2^64 p, so the
    // REDC quotient is below (5p / 2^64 + 1) p < 2.25 p.
    struct Acc {
        uint64_t hi, lo;
    };
    __device__ Acc prod(V a, V b) const
    {
        Acc t;
        mul128(a, b, t.hi, t.lo);
        return t;
    }
    __device__ Acc fma(Acc t, V a, V b) const
    {
        mac128(t.hi, t.lo, a, b);
        return t;
    }
    __device__ V reduce(Acc t) const
    {
        const uint64_t mm = t.lo * pinv;
        uint64_t u = t.hi + mulhi64(mm, p) + (t.lo != 0 ? 1 : 0);
        u = u >= p ? u - p : u;
        return u >= p ? u - p : u;
    }
    // pv x - fr q as pv x + fr (p - q) < 2 p^2: one reduction, quotient < 1.5 p
    __device__ V msub(V pv, V x, V fr, V q) const
    {
        const Acc t = fma(prod(pv, x), fr, p - q);
        return redc(t.hi, t.lo);
    }
    // K independent msubs sharing pv and fr, formed stage by stage so that
    // their dependent multiplies interleave
    template <int K>
    __device__ void msub_n(V pv, const V (&x)[K], V fr, const V (&q)[K], V (&out)[K]) const
    {
        Acc t[K];
#pragma unroll
        for (int i = 0; i < K; ++i)
            t[i] = fma(prod(pv, x[i]), fr, p - q[i]);
        uint64_t mm[K];
#pragma unroll
        for (int i = 0; i < K; ++i)
            mm[i] = t[i].lo * pinv;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const uint64_t u = t[i].hi + mulhi64(mm[i], p) + (t[i].lo != 0 ? 1 : 0);
            out[i] = u >= p ? u - p : u;
        }
    }
    __device__ V sub(V a, V b) const { return a >= b ? a - b : a + (p - b); }
    __device__ V add(V a, V b) const
    {
        const uint64_t t = a + b;  // < 2p < 2^63
        return t >= p ? t - p : t;
    }
    __device__ V one() const { return r1; }
    __device__ V to(uint64_t x) const { return mul(x, r2); }
    __device__ uint64_t from(V x) const { return redc(0, x); }
    static __device__ V load(const uint64_t *w) { return *w; }
    // a = x R: the plain inverse of a is x^-1 R^-1, and x^-1 R = mul(mul(a^-1, R^2), R^2)
    __device__ V inv(V a) const { return mul(mul(inv_euclid(a, p), r2), r2); }
};

struct Mont32Arith {
    using V = uint32_t;
    uint32_t p, pinv, r1, r2;  // pinv = -p^-1 mod 2^32, r1 = 2^32 mod p, r2 = 2^64 mod p
    __device__ explicit Mont32Arith(const ModConst &c)
        : p((uint32_t)c.p), pinv((uint32_t)c.pinvneg), r1((uint32_t)c.c32), r2((uint32_t)c.c64) {}
    __device__ V redc(uint64_t t) const
    {
        // t < p^2 < 2^60: t + m p < 2^60 + 2^62, and the quotient u is below 2p;
        // min(u, u - p) picks u - p exactly when u >= p (u - p wraps otherwise)
        const uint32_t mm = (uint32_t)t * pinv;
        const uint32_t u = (uint32_t)((t + (uint64_t)mm * p) >> 32);
        const uint32_t v = u - p;
        return v < u ? v : u;
    }
    __device__ V mul(V a, V b) const { return redc((uint64_t)a * b); }
    // Sums of up to five products, reduced once: t < 5 p^2 < 5 2^60, t + mm p
    // < 9 2^60 < 2^64 and the quotient is below (5p / 2^32 + 1) p < 2.25 p.
    using Acc = uint64_t;
    __device__ Acc prod(V a, V b) const { return (uint64_t)a * b; }
    __device__ Acc fma(Acc t, V a, V b) const { return t + (uint64_t)a * b; }
    __device__ V reduce(Acc t) const
    {
        const uint32_t mm = (uint32_t)t * pinv;
        uint32_t u = (uint32_t)((t + (uint64_t)mm * p) >> 32);
        uint32_t v = u - p;
        u = v < u ? v : u;
        v = u - p;
        return v < u ? v : u;
    }
    // pv x - fr q as pv x + fr (p - q) < 2 p^2: one reduction, quotient < 2p
    __device__ V msub(V pv, V x, V fr, V q) const
    {
        return redc((uint64_t)pv * x + (uint64_t)fr * (p - q));
    }
    // K independent msubs sharing pv and fr, formed stage by stage so that
    // their dependent multiplies interleave
    template <int K>
    __device__ void msub_n(V pv, const V (&x)[K], V fr, const V (&q)[K], V (&out)[K]) const
    {
        uint64_t t[K];
#pragma unroll
        for (int i = 0; i < K; ++i)
            t[i] = (uint64_t)pv * x[i] + (uint64_t)fr * (p - q[i]);
        uint32_t mm[K];
#pragma unroll
        for (int i = 0; i < K; ++i)
            mm[i] = (uint32_t)t[i] * pinv;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const uint32_t u = (uint32_t)((t[i] + (uint64_t)mm[i] * p) >> 32);
            const uint32_t v = u - p;
            out[i] = v < u ? v : u;
        }
    }
    __device__ V sub(V a, V b) const
    {
        // a - b, plus p when it wraps: min(d, d + p) as unsigned words
        const uint32_t d = a - b, e = d + p;
        return e < d ? e : d;
    }
    __device__ V add(V a, V b) const
    {
        // a + b < 2p < 2^31: min(s, s - p) as unsigned words
        const uint32_t t = a + b, u = t - p;
        return u < t ? u : t;
    }
    __device__ V one() const { return r1; }
    __device__ V to(uint64_t x) const { return mul((V)x, r2); }
    // a residue below p < 2^30 is the low half of its 64-bit word
    static __device__ V load(const uint64_t *w) { return *reinterpret_cast<const uint32_t *>(w); }
    __device__ uint64_t from(V x) const { return redc(x); }
    __device__ V inv(V a) const
    {
        // Fermat a^(p-2), in the Montgomery domain (r1 is 1 there)
        uint32_t e = p - 2;
        V r = r1, b = a;
        while (e) {
            if (e & 1)
                r = mul(r, b);
            b = mul(b, b);
            e >>= 1;
        }
        return r;
    }
};

// Moduli from 2^62 to 2^64, and p = 2: plain residues, every product
// reduced at once by the division of mod128_any (Moller-Granlund); sums of
// residues may carry out of 64 bits, so additions go through addmod_any.
// Slower than the Montgomery forms, which need p < 2^62 for their lazy sums.
struct BigArith {
    using V = uint64_t;
    ModConst mc;
    __device__ explicit BigArith(const ModConst &c) : mc(c) {}
    __device__ V mul(V a, V b) const { return mulmod_any(a, b, mc); }
    using Acc = uint64_t;  // a reduced running sum
    __device__ Acc prod(V a, V b) const { return mul(a, b); }
    __device__ Acc fma(Acc t, V a, V b) const { return addmod_any(t, mul(a, b), mc.p); }
    __device__ V reduce(Acc t) const { return t; }
    // pv x - fr q
    __device__ V msub(V pv, V x, V fr, V q) const
    {
        return addmod_any(mul(pv, x), mul(fr, q ? mc.p - q : 0), mc.p);
    }
    template <int K>
    __device__ void msub_n(V pv, const V (&x)[K], V fr, const V (&q)[K], V (&out)[K]) const
    {
#pragma unroll
        for (int i = 0; i < K; ++i)
            out[i] = msub(pv, x[i], fr, q[i]);
    }
    __device__ V sub(V a, V b) const { return a >= b ? a - b : a + (mc.p - b); }
    __device__ V add(V a, V b) const { return addmod_any(a, b, mc.p); }
    __device__ V one() const { return 1; }
    __device__ V to(uint64_t x) const { return x; }
    __device__ uint64_t from(V x) const { return x; }
    static __device__ V load(const uint64_t *w) { return *w; }
    __device__ V inv(V a) const
    {
        // Fermat a^(p-2)
        uint64_t e = mc.p - 2;
        V r = 1, b = a;
        while (e) {
            if (e & 1)
                r = mul(r, b);
            b = mul(b, b);
            e >>= 1;
        }
        return r;
    }
};

// workgroup barrier ordering LDS only: global loads in flight stay in flight
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// value of v in lane l (l uniform across the wave)
__device__ __forceinline__ uint32_t lane_value(uint32_t v, int l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ uint64_t lane_value(uint64_t v, int l)
{
    return ((uint64_t)lane_value((uint32_t)(v >> 32), l) << 32) | lane_value((uint32_t)v, l);
}

// ---------------------------------------------------------------------------
// kernel 1: permutation, left nullspace, step description
// ---------------------------------------------------------------------------
// Fraction-free Gauss-Jordan on At = (permuted residual)^T (n x m): pivot rows
// stay where they are (prow[k] is the row of the k-th pivot, pivcol[k] its
// column); for a pivot at (pr, c) every other row r becomes
// pv * row_r - At[r][c] * row_pr with pv = At[pr][c].  The k-th row of the
// reduced row echelon form is row prow[k] divided by its pivot; the step
// multiplies by the other pivots' product (the pivot's cofactor) instead, so
// its kernel vectors are the normalised ones times the product of all pivots.
//
// The columns are taken P at a time (a panel).  One wave runs the pivot
// steps of the panel on the panel's columns only and composes them: every row
// ends as C[r][0] * (its own row at the panel's start) + sum_j C[r][j+1] *
// (the pivot row of panel column j at the panel's start), the same exact
// residues as the one-column-at-a-time sequence.  The waves then apply that
// combination to their own columns while the next panel is composed: one
// barrier per panel instead of one per column.
template <typename Arith, int H>
__global__ void __launch_bounds__(kStepThreads) mbasis_step_kernel(LeafArgs a, int ord)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    using V = typename Arith::V;
    // panel width: 4 columns for 32-bit words, 2 for 64-bit ones in the
    // one-row kernels, whose dearer products make a shorter composition pay
    // for the extra barriers; the two-row kernels (more than 64 residual
    // columns) are bound by the apply, which wider panels make shorter
    constexpr int P = sizeof(V) == 8 && H == 1 ? 2 : kPanel;
    static_assert(P <= kPanel, "LDS is sized for kPanel");
    const int m = a.m, n = a.n;
    const Arith ar(a.mc);
    const int ms = m | 1;                   // odd row stride of At
    // 64-bit words first (sh), then the V arrays, then the int arrays
    // (layout and sizes as step_lds_bytes, whose Ck is sized for kPanel)
    int64_t *sh = reinterpret_cast<int64_t *>(smem);                 // m
    V *At = reinterpret_cast<V *>(sh + m);                           // n x ms
    V *cof = At + n * ms;                                            // kMaxN: pivot cofactors
    V *Ck = cof + kMaxN;                                             // 2 x kMaxN x (P + 1)
    int *perm = reinterpret_cast<int *>(Ck + 2 * kMaxN * (kPanel + 1));  // m
    int *plist = perm + m;                                           // m
    int *colpiv = plist + m;                                         // m (columns of At)
    int *pos = colpiv + m;                                           // m (then: pivot row by column)
    int *pivcol = pos + m;                                           // kMaxN (by rank)
    int *prow = pivcol + kMaxN;                                      // kMaxN (by rank)
    int *ispiv = prow + kMaxN;                                       // kMaxN (rows of At)
    __shared__ int s_nb[2], s_prs[2][kPanel];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint64_t p = a.mc.p;

    // Shifts, control words and this step's residual coefficient.  The
    // coefficient sits in the fixed buffer Rn (m x n, row-major), written by
    // the previous update or the leaf's initialisation, so its loads wait for
    // no control word and are contiguous across the workgroup; the shifts are
    // stored to LDS before those loads are issued (a wait on a vector load
    // also waits on those issued before it).  Thread 0 writes the control
    // words back only after full barriers.
    const int64_t shv = tid < m ? a.shift[tid] : 0;  // m <= kMaxM <= kStepThreads
    int *c = a.ctl;
    const int flip = c[C_FLIP], cur0 = c[C_CUR], done = c[C_DONE];
    if (tid < m) {
        sh[tid] = shv;
        pos[tid] = 0;
        colpiv[tid] = 0;
    }
    const int cur = flip ? cur0 ^ 1 : cur0;
    // element q = tid + k kStepThreads of the coefficient (row q / n of the
    // unpermuted residual, column q % n), stored to At[q % n][pos[q / n]] once
    // the permutation is known; the first kPre per thread are loaded before
    // the sort (indices past m n clamped, their values unused)
    const int mn = m * n;
    constexpr int kPre = 4;
    V pre[kPre];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int q = tid + k * kStepThreads;
        pre[k] = Arith::load(a.Rn + (q < mn ? q : mn - 1));
    }
    for (int r = tid; r < kMaxN; r += kStepThreads)
        ispiv[r] = 0;
    if (done) {
        __syncthreads();
        if (tid == 0) {
            const int len = c[C_LEN] + c[C_GROW];
            c[C_CUR] = cur;
            c[C_FLIP] = 0;
            c[C_LEN] = len;
            c[C_GROW] = 0;
            c[C_MODE] = 0;
        }
        return;
    }
    lds_barrier();

    // stable permutation sorting (shift, index): pos[i] = #{j : (sh_j, j) < (sh_i, i)}
    for (int idx = tid; idx < m * m; idx += kStepThreads) {
        const int j = idx / m, i = idx - j * m;  // neighbouring lanes update different counters
        const int64_t si = sh[i], sj = sh[j];
        if ((sj < si) || (sj == si && j < i))
            atomicAdd(&pos[i], 1);
    }
    lds_barrier();
    for (int i = tid; i < m; i += kStepThreads)
        perm[pos[i]] = i;
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int q = tid + k * kStepThreads;
        if (q < mn) {
            const int i = q / n, j = q - i * n;
            At[j * ms + pos[i]] = ar.to(pre[k]);
        }
    }
    for (int q = tid + kPre * kStepThreads; q < mn; q += kStepThreads) {
        const int i = q / n, j = q - i * n;
        At[j * ms + pos[i]] = ar.to(Arith::load(a.Rn + q));
    }
    __syncthreads();

    // Gauss-Jordan by panels of P columns.  Wave w owns the columns
    // [w cw, (w + 1) cw) of At and applies every panel's combination to them;
    // the owner of the next panel then composes it from its own updated
    // columns (the LDS operations of one wave take effect in program order)
    // while the other waves are still applying: one barrier per panel, with
    // the coefficients double-buffered.
    int rank = 0;
    const int wave = tid >> 6;
    // columns per wave: P times a power of two (so that cw divides the
    // wave's 64 lanes), 16 waves covering m
    int cw = P;
    while (16 * cw < m)
        cw *= 2;
    // compose the panel of columns c0 .. c0 + P - 1 into buffer buf (the
    // lambdas are forced inline: called from two places, compose was left as
    // a call in the two-row kernels, its captures and arrays on the stack)
    auto compose = [&](const int c0, const int buf) __attribute__((always_inline)) {
        __builtin_amdgcn_s_setprio(2);
        {
            // rows lane + 64 h (h < H) of the panel, their pivot flags and the
            // coefficients C[.][0..P] of the composition
            V pn[H][P], cc[H][P + 1];
            bool isp[H], ok[H];
            int prs[P] = {0};
            int tk = 0;  // panel columns that took a pivot
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int r = lane + 64 * h;
                ok[h] = r < n;
                isp[h] = ok[h] && ispiv[r];
#pragma unroll
                for (int j = 0; j < P; ++j)
                    pn[h][j] = ok[h] && c0 + j < m ? At[r * ms + c0 + j] : (V)0;
                cc[h][0] = ar.one();
#pragma unroll
                for (int s2 = 1; s2 <= P; ++s2)
                    cc[h][s2] = 0;
            }
            int nb = 0;
            // one pivot step on panel column j (a compile-time index: the
            // convergent ballots and lane reads keep the compiler from unrolling
            // a loop around them, and the arrays must stay in registers)
            auto step = [&](auto jc) __attribute__((always_inline)) {
                constexpr int j = decltype(jc)::value;
                // columns past m, or a full rank already reached, take no pivot
                const bool open = c0 + j < m && rank + nb < n;
                const unsigned long long b0 = __ballot(open && ok[0] && !isp[0] && pn[0][j] != 0);
                const unsigned long long b1 =
                    H == 2 ? __ballot(open && ok[H - 1] && !isp[H - 1] && pn[H - 1][j] != 0) : 0ull;
                if (!b0 && !b1)
                    return;  // no pivot in this column
                const int pr = b0 ? __ffsll(b0) - 1 : 63 + __ffsll(b1);
                const int pl = pr & 63, ph = pr >> 6;
                // broadcast the pivot row (panel values and coefficients): the
                // pivot lane is uniform, so its registers are read into
                // scalar registers; its self coefficient cc[0] moves to slot
                // j + 1
                V qp[P], qc[P + 1];
#pragma unroll
                for (int jj = j; jj < P; ++jj)
                    qp[jj] = lane_value(H == 2 && ph ? pn[H - 1][jj] : pn[0][jj], pl);
#pragma unroll
                for (int s2 = 1; s2 <= j; ++s2)
                    qc[s2] = lane_value(H == 2 && ph ? cc[H - 1][s2] : cc[0][s2], pl);
                qc[j + 1] = lane_value(H == 2 && ph ? cc[H - 1][0] : cc[0][0], pl);
                const V pv = qp[j];
                // Only the nonzero terms are formed: column j of every other
                // row becomes pv * fr - fr * pv = 0, and slots past j are zero
                // in every row until this step, so slot j + 1 is -fr * qc.
                // Every lane computes and the pivot lane keeps its own values
                // (selects, not branches, so the step stays one basic block);
                // rows past n hold zeros and are never published.
#pragma unroll
                for (int h = 0; h < H; ++h) {
                    const bool me = lane + 64 * h == pr;
                    const V fr = pn[h][j];
                    // the P products of this step: panel columns j + 1 .. P - 1,
                    // slots 1 .. j, and slot j + 1 (from zero)
                    V xs[P], qs[P], ts[P];
#pragma unroll
                    for (int jj = j + 1; jj < P; ++jj) {
                        xs[jj - j - 1] = pn[h][jj];
                        qs[jj - j - 1] = qp[jj];
                    }
#pragma unroll
                    for (int s2 = 1; s2 <= j; ++s2) {
                        xs[P - j - 2 + s2] = cc[h][s2];
                        qs[P - j - 2 + s2] = qc[s2];
                    }
                    xs[P - 1] = 0;
                    qs[P - 1] = qc[j + 1];
                    ar.msub_n(pv, xs, fr, qs, ts);
#pragma unroll
                    for (int jj = j + 1; jj < P; ++jj)
                        pn[h][jj] = me ? pn[h][jj] : ts[jj - j - 1];
                    pn[h][j] = me ? fr : (V)0;
#pragma unroll
                    for (int s2 = 1; s2 <= j; ++s2)
                        cc[h][s2] = me ? cc[h][s2] : ts[P - j - 2 + s2];
                    cc[h][j + 1] = me ? qc[j + 1] : ts[P - 1];
                    const V t0 = ar.mul(pv, cc[h][0]);
                    cc[h][0] = me ? (V)0 : t0;
                    isp[h] = isp[h] || me;
                }
                prs[j] = pr;
                tk |= 1 << j;
                ++nb;
            };
            static_assert(P >= 1 && P <= 4, "the panel steps below are written out for up to 4 columns");
            step(std::integral_constant<int, 0>{});
            if constexpr (P > 1)
                step(std::integral_constant<int, 1>{});
            if constexpr (P > 2)
                step(std::integral_constant<int, 2>{});
            if constexpr (P > 3)
                step(std::integral_constant<int, 3>{});
            // publish: coefficients, pivot flags, pivot rows
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int r = lane + 64 * h;
                if (ok[h]) {
#pragma unroll
                    for (int s2 = 0; s2 <= P; ++s2)
                        Ck[(buf * kMaxN + r) * (P + 1) + s2] = cc[h][s2];
                    ispiv[r] = isp[h] ? 1 : 0;
                }
            }
            if (lane == 0) {
                // pivot bookkeeping, read only after the elimination
                int q = rank;
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    if (tk >> k & 1) {
                        prow[q] = prs[k];
                        pivcol[q] = c0 + k;
                        colpiv[c0 + k] = 1;
                        pos[c0 + k] = prs[k];  // the column's pivot row
                        ++q;
                    }
                }
                s_nb[buf] = nb;
#pragma unroll
                for (int k = 0; k < P; ++k)
                    s_prs[buf][k] = prs[k];
            }
        }
        __builtin_amdgcn_s_setprio(0);
    };
    // Residuals of at most 32 columns (n <= 32 rows of At) leave the upper
    // half of the composing wave idle in compose().  compose_half() puts the
    // coefficient slots there: upper lane 32 + r holds slots 0, 1, 2 of row r
    // (slots 0, 1 for P = 2) and lower lane r the row's panel values and the
    // last slots, in registers whose lower halves pass from a panel column to
    // a slot once that column is eliminated.  Each instruction then advances
    // a panel value and a coefficient at once: three multiply-subtracts per
    // pivot step instead of five (two instead of three for P = 2).
    //   P = 4, registers Z A B C: lower lane r holds row r in panel columns
    //   0..3, A becomes slot 3 before step 2 and B slot 4 before step 3;
    //   upper lane 32 + r holds slots 0, 1, 2 in A, B, C.
    //   P = 2, registers Z A B: lower lane r holds columns 0, 1 in Z, A and
    //   slot 2 in B; upper lane 32 + r holds slots 0, 1 in A, B.
    // Slot 0 takes pv * c0 as the multiply-subtract with q = 0 (pv c0 + fr p,
    // the same residue); a register that starts a slot holds zero, so the
    // slot's first value is -fr * q as in compose().
    auto compose_half = [&](const int c0, const int buf) __attribute__((always_inline)) {
        __builtin_amdgcn_s_setprio(2);
        {
            static_assert(P == 2 || P == 4, "compose_half is written out for panels of 2 and 4");
            const bool hi = lane >= 32;
            const int r = lane & 31;
            const bool ok = r < n;
            auto col = [&](const int j) -> V {
                return !hi && ok && c0 + j < m ? At[r * ms + c0 + j] : (V)0;
            };
            V Z = col(0);
            V A = hi ? ar.one() : col(1);
            V B = P == 4 ? col(2) : (V)0;
            V C = P == 4 ? col(3) : (V)0;
            bool isp = !hi && ok && ispiv[r];
            int prs[P] = {0};
            int tk = 0, nb = 0;
            // every lane gets the value of x in lane (lane & 31)
            auto low = [](const V x) -> V {
                if constexpr (sizeof(V) == 4) {
                    return __builtin_amdgcn_permlane32_swap(x, x, false, false)[0];
                } else {
                    const uint32_t l = (uint32_t)x, h = (uint32_t)(x >> 32);
                    const uint32_t l2 = __builtin_amdgcn_permlane32_swap(l, l, false, false)[0];
                    const uint32_t h2 = __builtin_amdgcn_permlane32_swap(h, h, false, false)[0];
                    return ((uint64_t)h2 << 32) | l2;
                }
            };
            auto step = [&](auto jc) __attribute__((always_inline)) {
                constexpr int j = decltype(jc)::value;
                if constexpr (P == 4 && j == 2)
                    A = hi ? A : (V)0;  // column 1 is done: lower A is slot 3
                if constexpr (P == 4 && j == 3)
                    B = hi ? B : (V)0;  // column 2 is done: lower B is slot 4
                const V xj = j == 0 ? Z : j == 1 ? A : j == 2 ? B : C;
                const bool open = c0 + j < m && rank + nb < n;
                const unsigned long long b = __ballot(open && !hi && ok && !isp && xj != 0);
                if (!b)
                    return;  // no pivot in this column
                const int pr = __ffsll(b) - 1;  // a lower lane
                const bool me = r == pr;
                const V pv = lane_value(xj, pr);
                const V fr = low(xj);
                const V Q = lane_value(A, 32 + pr);  // the pivot row's slot 0
                constexpr int G = P == 4 ? 3 : 2;
                V xs[G], qs[G], ts[G], kp[G];  // operands, and what the pivot row keeps
                const V zero = 0;
                if constexpr (P == 4) {
                    xs[0] = A;
                    xs[1] = B;
                    xs[2] = C;
                    if constexpr (j == 0) {
                        const V p1 = lane_value(A, pr), p2 = lane_value(B, pr), p3 = lane_value(C, pr);
                        qs[0] = hi ? zero : p1;
                        qs[1] = hi ? Q : p2;
                        qs[2] = hi ? zero : p3;
                        kp[0] = hi ? zero : A;
                        kp[1] = hi ? Q : B;
                        kp[2] = C;
                    } else if constexpr (j == 1) {
                        const V p2 = lane_value(B, pr), p3 = lane_value(C, pr), q1 = lane_value(B, 32 + pr);
                        qs[0] = zero;
                        qs[1] = hi ? q1 : p2;
                        qs[2] = hi ? Q : p3;
                        kp[0] = hi ? zero : A;
                        kp[1] = B;
                        kp[2] = hi ? Q : C;
                    } else if constexpr (j == 2) {
                        const V p3 = lane_value(C, pr), q1 = lane_value(B, 32 + pr), q2 = lane_value(C, 32 + pr);
                        qs[0] = hi ? zero : Q;
                        qs[1] = hi ? q1 : zero;
                        qs[2] = hi ? q2 : p3;
                        kp[0] = hi ? zero : Q;
                        kp[1] = B;
                        kp[2] = C;
                    } else {
                        const V q3 = lane_value(A, pr), q1 = lane_value(B, 32 + pr), q2 = lane_value(C, 32 + pr);
                        qs[0] = hi ? zero : q3;
                        qs[1] = hi ? q1 : Q;
                        qs[2] = hi ? q2 : zero;
                        kp[0] = hi ? zero : A;
                        kp[1] = hi ? B : Q;
                        kp[2] = C;
                    }
                } else {
                    xs[0] = A;
                    xs[1] = B;
                    if constexpr (j == 0) {
                        const V p1 = lane_value(A, pr);
                        qs[0] = hi ? zero : p1;
                        qs[1] = hi ? Q : zero;
                        kp[0] = hi ? zero : A;
                        kp[1] = hi ? Q : B;
                    } else {
                        const V q1 = lane_value(B, 32 + pr);
                        qs[0] = zero;
                        qs[1] = hi ? q1 : Q;
                        kp[0] = hi ? zero : A;
                        kp[1] = hi ? B : Q;
                    }
                }
                ar.msub_n(pv, xs, fr, qs, ts);
                A = me ? kp[0] : ts[0];
                B = me ? kp[1] : ts[1];
                if constexpr (P == 4)
                    C = me ? kp[2] : ts[2];
                isp = isp || me;  // read in the lower half only
                prs[j] = pr;
                tk |= 1 << j;
                ++nb;
            };
            step(std::integral_constant<int, 0>{});
            step(std::integral_constant<int, 1>{});
            if constexpr (P == 4) {
                step(std::integral_constant<int, 2>{});
                step(std::integral_constant<int, 3>{});
            }
            // publish: coefficients, pivot flags, pivot rows
            if (ok) {
                V *ck = Ck + (buf * kMaxN + r) * (P + 1);
                if (hi) {
                    ck[0] = A;
                    ck[1] = B;
                    if constexpr (P == 4)
                        ck[2] = C;
                } else {
                    if constexpr (P == 4) {
                        ck[3] = A;
                        ck[4] = B;
                    } else {
                        ck[2] = B;
                    }
                    ispiv[r] = isp ? 1 : 0;
                }
            }
            if (lane == 0) {
                // pivot bookkeeping, read only after the elimination
                int q = rank;
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    if (tk >> k & 1) {
                        prow[q] = prs[k];
                        pivcol[q] = c0 + k;
                        colpiv[c0 + k] = 1;
                        pos[c0 + k] = prs[k];  // the column's pivot row
                        ++q;
                    }
                }
                s_nb[buf] = nb;
#pragma unroll
                for (int k = 0; k < P; ++k)
                    s_prs[buf][k] = prs[k];
            }
        }
        __builtin_amdgcn_s_setprio(0);
    };
    // the split composition needs one row per lane and the Montgomery forms
    constexpr bool kHalf = H == 1 && !std::is_same<Arith, BigArith>::value;
    const bool half = kHalf && n <= 32;
    auto compose_panel = [&](const int c0, const int buf) __attribute__((always_inline)) {
        if constexpr (kHalf) {
            if (half) {
                compose_half(c0, buf);
                return;
            }
        }
        compose(c0, buf);
    };
    // apply the combination in buffer buf to this wave's columns: the pivot
    // rows' values are read before any row is written.  Every term is
    // formed: a panel column without a pivot has row 0 in s_prs and zero
    // coefficients in every row, so its term is zero, and issuing all loads
    // before the products (no branch between them) keeps the LDS round
    // trips of a panel to three instead of one per term.
    // A pivot column of an earlier panel (below cdone) is zero but in its
    // pivot row, and no later pivot row has a nonzero there: the combination
    // only multiplies that row's value by the row's own coefficient.
    const int acl = lane & (cw - 1), ag = lane / cw, aG = 64 / cw;  // cw: a power of two
    const int acg = wave * cw + acl;
    auto apply = [&](const int buf, const int cdone) __attribute__((always_inline)) {
        if (acg >= m)
            return;
        if (acg < cdone && colpiv[acg]) {
            const int r = pos[acg];
            if (ag == (r & (aG - 1))) {
                V *x = At + r * ms + acg;
                *x = ar.mul(Ck[(buf * kMaxN + r) * (P + 1)], *x);
            }
            return;
        }
        int prq[P];
#pragma unroll
        for (int q = 0; q < P; ++q)
            prq[q] = s_prs[buf][q];
        V pvr[P];
#pragma unroll
        for (int q = 0; q < P; ++q)
            pvr[q] = At[prq[q] * ms + acg];
        const V *C = Ck + buf * kMaxN * (P + 1);
        for (int r = ag; r < n; r += aG) {
            const V *cr = C + r * (P + 1);
            V *x = At + r * ms + acg;
            V cv[P + 1];
#pragma unroll
            for (int s2 = 0; s2 <= P; ++s2)
                cv[s2] = cr[s2];
            typename Arith::Acc acc = ar.prod(cv[0], *x);
#pragma unroll
            for (int q = 0; q < P; ++q)
                acc = ar.fma(acc, cv[q + 1], pvr[q]);
            *x = ar.reduce(acc);
        }
    };
    if (wave == 0)
        compose_panel(0, 0);
    __syncthreads();
    for (int k = 0;; ++k) {
        const int buf = k & 1;
        const int nb = s_nb[buf];
        if (nb > 0)
            apply(buf, k * P);
        rank += nb;
        const int c1 = (k + 1) * P;
        if (c1 >= m || rank >= n)
            break;
        if (wave == c1 / cw)
            compose_panel(c1, buf ^ 1);
        __syncthreads();
    }
    __syncthreads();
    // Cofactors of the pivots of the reduced echelon form: for pivot k, the
    // product of all the others (Montgomery's simultaneous-inversion trick
    // without the inversion): the kernel vectors are then D times the
    // normalised ones, D the product of all pivots, and no inverse is needed
    // in the step.  Wave 0, lane l holding pivots l and l + 64 (rank <= 128):
    // prefix and suffix products by scans across the wave.  Pivots are
    // nonzero, so is D.
    __shared__ V s_D;
    if (wave == 0) {
        const V x0 = lane < rank ? At[prow[lane] * ms + pivcol[lane]] : ar.one();
        const V x1 = lane + 64 < rank ? At[prow[lane + 64] * ms + pivcol[lane + 64]] : ar.one();
        const V v = ar.mul(x0, x1);
        V pre = v, suf = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const V up = __shfl_up(pre, d);
            const V dn = __shfl_down(suf, d);
            if (lane >= d)
                pre = ar.mul(pre, up);
            if (lane + d < 64)
                suf = ar.mul(suf, dn);
        }
        V ex = __shfl_up(pre, 1), exs = __shfl_down(suf, 1);
        if (lane == 0)
            ex = ar.one();
        if (lane == 63)
            exs = ar.one();
        const V others = ar.mul(ex, exs);  // every other lane's pair
        if (lane < rank)
            cof[lane] = ar.mul(others, x1);
        if (lane + 64 < rank)
            cof[lane + 64] = ar.mul(others, x0);
        if (lane == 63)
            s_D = pre;
    }
    // permutation list of left_nullspace_compact: pivot columns, then the
    // others; the last wave (the cofactors run on the first one) counts the
    // pivot columns before each column by ballots
    if (tid >= kStepThreads - 64) {
        const int l = tid - (kStepThreads - 64);
        int base = 0;
#pragma unroll
        for (int h = 0; h < kMaxM / 64; ++h) {
            const int i = l + 64 * h;
            const bool pc = i < m && colpiv[i];
            const unsigned long long b = __ballot(pc);
            const int before = base + __popcll(b & ((1ull << l) - 1));
            if (i < m) {
                if (pc)
                    plist[before] = i;
                else
                    plist[rank + i - before] = i;
            }
            base += __popcll(b);
        }
    }
    __syncthreads();
    const int nullity = m - rank;
    const int mode = nullity == 0 ? 2 : (nullity == m ? 0 : 1);
    if (mode == 1) {
        for (int idx = tid; idx < nullity * rank; idx += kStepThreads) {
            const int t = idx / rank, r = idx - t * rank;
            const uint64_t v = ar.from(ar.mul(At[prow[r] * ms + plist[rank + t]], cof[r]));
            a.X[(size_t)t * m + r] = v ? p - v : 0;
        }
        for (int i = tid; i < m; i += kStepThreads) {
            const int row = perm[plist[i]];
            a.piv[i] = row;
            a.rowinfo[row] = i < rank ? -1 - i : i - rank;
            if (i < rank)
                a.shift[row] = sh[row] + 1;
            // rows updated by a kernel vector take the factor D; the scales
            // start at 1 (Montgomery form) at the leaf's first step
            if (ord == 0 || i >= rank) {
                const V s0 = ord == 0 ? ar.one() : (V)a.scale[row];
                a.scale[row] = i >= rank ? ar.mul(s0, s_D) : s0;
            }
        }
        if (tid == 0)
            a.scale[m] = sizeof(V) == 8 ? ar.from(s_D) : (uint64_t)s_D;
    } else if (mode == 2) {
        for (int i = tid; i < m; i += kStepThreads)
            a.shift[i] = sh[i] + (a.ordL - ord);
    }
    if (mode != 1 && ord == 0)
        for (int i = tid; i < m; i += kStepThreads)
            a.scale[i] = ar.one();
    if (tid == 0) {
        const int len = c[C_LEN] + c[C_GROW];
        c[C_CUR] = cur;
        c[C_GROW] = 0;
        c[C_LEN] = len;
        c[C_MODE] = mode;
        c[C_RANK] = rank;
        c[C_FLIP] = mode != 0 ? 1 : 0;
        if (mode == 2) {
            c[C_SHIFT] = a.ordL - ord;
            c[C_LOLD] = len;
            c[C_LEN] = len + (a.ordL - ord);
            c[C_DONE] = 1;
        }
    }
}

// ---------------------------------------------------------------------------
// kernel 2: apply the step to appbas (and to the residual series)
// ---------------------------------------------------------------------------
// base + sum_r X[t][r] * src[row piv[r], column j, coefficient d].  All
// loads of a group of terms (32 words below 2^30, or 16 64-bit words: one L2
// round trip) are issued before its products, with no branch between them:
// terms past the rank read row piv[r] of the LDS row list (padded with row 0
// up to kMaxM; rank <= kMaxN) and X words past the row (X carries kXPad words
// of padding),
// and their X coefficient is zeroed by a select after the load.
template <bool SMALL, bool BIG>
__device__ __forceinline__ uint64_t row_combo(const uint64_t *src, const LeafArgs &a, const int *piv, int t,
                                              int rank, int stride, int cols, int j, int d, uint64_t base)
{
    const uint64_t *xr = a.X + (size_t)t * a.m;
    const uint64_t *sj = src + (size_t)j * stride + d;
    const size_t rstride = (size_t)cols * stride;
    constexpr int G = SMALL ? 32 : 16;
    uint64_t acc = base;
    for (int r = 0; r < rank; r += G) {
        if constexpr (SMALL) {
            // p < 2^30: words are read through their low halves; 16 products
            // stay below 16 (p - 1)^2 < 2^64 - 2^35, so each half-group is
            // summed with a residue before one Barrett reduction
            uint32_t sv[G], xv[G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                sv[u] = (uint32_t)sj[(size_t)piv[r + u] * rstride];
                const uint32_t x = (uint32_t)xr[r + u];
                xv[u] = r + u < rank ? x : 0u;
            }
#pragma unroll
            for (int h = 0; h < G; h += 16) {
                uint64_t part = 0;
#pragma unroll
                for (int u = h; u < h + 16; ++u)
                    part += (uint64_t)xv[u] * sv[u];
                acc = barrett(acc + part, a.mc.bm, a.mc.p);
            }
        } else if constexpr (BIG) {
            // p >= 2^62: every product reduced at once, the sum kept below p
            uint64_t sv[G], xv[G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                sv[u] = sj[(size_t)piv[r + u] * rstride];
                const uint64_t x = xr[r + u];
                xv[u] = r + u < rank ? x : 0;
            }
#pragma unroll
            for (int u = 0; u < G; ++u)
                acc = addmod_any(acc, mulmod_any(xv[u], sv[u], a.mc), a.mc.p);
        } else {
            // 64-bit words (p < 2^62): an exact 128-bit sum of 16 products
            // and a residue, 16 (p - 1)^2 + p < 2^128 - 2^67, folded once (the
            // 128-bit products are the bound here, not the loads)
            uint64_t sv[G], xv[G];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                sv[u] = sj[(size_t)piv[r + u] * rstride];
                const uint64_t x = xr[r + u];
                xv[u] = r + u < rank ? x : 0;
            }
            uint64_t hi = 0, lo = acc;
#pragma unroll
            for (int u = 0; u < G; ++u)
                mac128(hi, lo, xv[u], sv[u]);
            acc = reduce128(hi, lo, a.mc);
        }
    }
    return acc;
}

template <bool SMALL, bool BIG>
__global__ void __launch_bounds__(256) mbasis_update_kernel(LeafArgs a, int ord)
{
    // The control words, the step's row list and row roles and its factor D
    // are all requested at entry, before any branch on them: one round trip
    // to memory instead of three dependent ones (mode, then the row lists,
    // then rank, length and D).  In modes 0 and 2 the lists and D are stale
    // and unused.
    static_assert(kMaxM == 256, "one row-list entry per thread of the 256-thread workgroup");
    const int mode = a.ctl[C_MODE];
    const int cur = a.ctl[C_CUR];
    const int rank = a.ctl[C_RANK];
    const int L = a.ctl[C_LEN];
    const uint64_t D = a.scale[a.m];
    const int tx = threadIdx.x;
    const int piv_t = tx < a.m ? a.piv[tx] : 0;  // groups read up to kMaxM entries
    const int rinfo_t = tx < a.m ? a.rowinfo[tx] : 0;
    const unsigned tstride = gridDim.x * blockDim.x;
    const unsigned t0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (mode == 0) {
        // residual unchanged: the coming step's coefficient is ord + 1
        // (nothing to do once the basis is complete)
        if (ord + 1 < a.ordL && !a.ctl[C_DONE]) {
            const uint64_t *rs = a.R[cur];
            for (unsigned e = t0; e < (unsigned)(a.m * a.n); e += tstride)
                a.Rn[e] = rs[(size_t)e * a.ordL + ord + 1];
        }
        return;
    }
    const int m = a.m, n = a.n, cap = a.cap;
    const uint64_t *src = a.app[cur];
    uint64_t *dst = a.app[cur ^ 1];
    if (mode == 2) {
        const int s = a.ctl[C_SHIFT], Lold = a.ctl[C_LOLD];
        const unsigned Lnew = (unsigned)(Lold + s);
        const unsigned items = (unsigned)(m * m) * Lnew;
        for (unsigned idx = t0; idx < items; idx += tstride) {
            const unsigned e = idx / Lnew;
            const int d = (int)(idx - e * Lnew);
            dst[(size_t)e * cap + d] = d >= s ? src[(size_t)e * cap + d - s] : 0;
        }
        return;
    }
    __shared__ int piv[kMaxM], rinfo[kMaxM];
    piv[tx] = piv_t;
    rinfo[tx] = rinfo_t;
    __syncthreads();
    // D: the kernel vectors' own coefficient (the step's pivot product)
    auto own = [&](uint64_t x) -> uint64_t {
        if constexpr (SMALL)  // D in Montgomery form: REDC(x D 2^32) = x D mod p
            return Mont32Arith(a.mc).redc((uint64_t)(uint32_t)x * (uint32_t)D);
        else
            return mulmod_g(x, D, a.mc);
    };
    const unsigned L1 = (unsigned)(L + 1 <= cap ? L + 1 : cap);
    const unsigned items = (unsigned)(m * m) * L1;
    // residual series: coefficients ord+1 .. ordL-1
    const uint64_t *rs = a.R[cur];
    uint64_t *rd = a.R[cur ^ 1];
    const int spn = a.ordL - ord - 1;
    const unsigned span = spn > 0 ? (unsigned)spn : 0u;
    const unsigned ritems = (unsigned)(m * n) * span;
    // one index space, basis items then residual items: with the grid sized
    // for both, a thread has one item
    for (unsigned idx = t0; idx < items + ritems; idx += tstride) {
        if (idx < items) {
            const unsigned e = idx / L1;
            const int d = (int)(idx - e * L1);
            const int i = (int)(e / (unsigned)m), j = (int)(e - (unsigned)i * m);
            const int ri = rinfo[i];
            uint64_t v;
            if (ri >= 0) {
                v = d < L ? row_combo<SMALL, BIG>(src, a, piv, ri, rank, cap, m, j, d, own(src[(size_t)e * cap + d]))
                          : 0;
            } else {
                v = d >= 1 ? src[(size_t)e * cap + d - 1] : 0;
                if (d == L && v != 0)
                    atomicOr(&a.ctl[C_GROW], 1);
            }
            dst[(size_t)e * cap + d] = v;
        } else {
            const unsigned k = idx - items;
            const unsigned e = k / span;
            const int d = ord + 1 + (int)(k - e * span);
            const int i = (int)(e / (unsigned)n), j = (int)(e - (unsigned)i * n);
            const int ri = rinfo[i];
            const uint64_t v =
                ri >= 0 ? row_combo<SMALL, BIG>(rs, a, piv, ri, rank, a.ordL, n, j, d, own(rs[(size_t)e * a.ordL + d]))
                        : rs[(size_t)e * a.ordL + d - 1];
            rd[(size_t)e * a.ordL + d] = v;
            if (d == ord + 1)
                a.Rn[e] = v;  // the coming step's coefficient
        }
    }
}

// scale[i] <- 1 / scale[i] as a plain residue (the scales are Montgomery
// residues, nonzero as products of pivots)
template <typename Arith>
__global__ void __launch_bounds__(kMaxM) leaf_scale_inv_kernel(LeafArgs a)
{
    const Arith ar(a.mc);
    const int i = threadIdx.x;
    if (i < a.m)
        a.scale[i] = ar.from(ar.inv((typename Arith::V)a.scale[i]));
}

// coefficients [0, len) of the m x m basis (entries of capacity cap) times
// their row's inverse scale
__global__ void __launch_bounds__(256) leaf_rescale_kernel(uint64_t *__restrict__ P, int64_t cap, int m,
                                                           int64_t len, const uint64_t *__restrict__ iscale,
                                                           ModConst mc)
{
    const int64_t items = (int64_t)m * m * len;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < items;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = idx / len, d = idx - e * len;
        const int row = (int)(e / m);
        P[e * cap + d] = mulmod_g(P[e * cap + d], iscale[row], mc);
    }
}

// appbas <- identity (capacity cap), R <- F truncated to ordL, control reset
__global__ void leaf_init_kernel(LeafArgs a, const uint64_t *__restrict__ F, int64_t strideF,
                                 int64_t lenF)
{
    const int m = a.m, n = a.n, cap = a.cap;
    const int64_t tstride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t itemsA = (int64_t)m * m * cap;
    for (int64_t idx = t0; idx < itemsA; idx += tstride) {
        const int64_t e = idx / cap;
        const int d = (int)(idx - e * cap);
        const int i = (int)(e / m), j = (int)(e - (int64_t)i * m);
        a.app[0][idx] = (d == 0 && i == j) ? 1 : 0;
    }
    const int64_t itemsR = (int64_t)m * n * a.ordL;
    for (int64_t idx = t0; idx < itemsR; idx += tstride) {
        const int64_t e = idx / a.ordL;
        const int d = (int)(idx - e * a.ordL);
        const uint64_t v = d < lenF ? F[e * strideF + d] : 0;
        a.R[0][idx] = v;
        if (d == 0)
            a.Rn[e] = v;
    }
    if (t0 < C_WORDS)
        a.ctl[t0] = (t0 == C_LEN) ? 1 : 0;
}

// copies an m x m x L block between two capacities (entry-major)
__global__ void copy_poly_kernel(uint64_t *__restrict__ dst, int64_t dcap,
                                 const uint64_t *__restrict__ src, int64_t scap, int64_t entries,
                                 int64_t len)
{
    const int64_t items = entries * len;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < items;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = idx / len, d = idx - e * len;
        dst[e * dcap + d] = src[e * scap + d];
    }
}

unsigned grid_for(int64_t items)
{
    int64_t b = (items + 255) / 256;
    if (b < 1)
        b = 1;
    if (b > 4096)
        b = 4096;
    return (unsigned)b;
}

// Bump allocator over one device buffer; results are moved down to the
// caller's mark so that temporaries of a recursion level are reclaimed.
struct Arena {
    char *base = nullptr;
    size_t size = 0, top = 0;
    uint64_t *alloc(size_t words)
    {
        size_t bytes = (words * 8 + 255) & ~(size_t)255;
        if (top + bytes > size)
            return nullptr;
        uint64_t *p = reinterpret_cast<uint64_t *>(base + top);
        top += bytes;
        return p;
    }
};

struct Poly {
    uint64_t *d;
    int64_t cap, len;
};

constexpr int kPmbasisThres = 32;  // PMBASIS_THRES, nmod_poly_mat_approximant.h:101

// M-Basis of F (m x n, entries at F + e*strideF, length lenF) at order ordL;
// the result is placed at the arena mark.
pmlg_status leaf(pmlg_ctx ctx, Arena &ar, const uint64_t *F, int64_t strideF, int64_t lenF,
                 int m, int n, int ordL, int64_t *dShift, Poly *out, hipStream_t s)
{
    const size_t mark = ar.top;
    const int cap = ordL + 1;
    LeafArgs a{};
    a.m = m;
    a.n = n;
    a.ordL = ordL;
    a.cap = cap;
    a.app[0] = ar.alloc((size_t)m * m * cap);
    a.app[1] = ar.alloc((size_t)m * m * cap);
    a.R[0] = ar.alloc((size_t)m * n * (ordL > 0 ? ordL : 1));
    a.R[1] = ar.alloc((size_t)m * n * (ordL > 0 ? ordL : 1));
    a.Rn = ar.alloc((size_t)m * n);
    a.scale = ar.alloc((size_t)m + 1);
    a.X = ar.alloc((size_t)m * m + kXPad);
    uint64_t *ints = ar.alloc((size_t)m + C_WORDS + 8);
    if (!a.app[0] || !a.app[1] || !a.R[0] || !a.R[1] || !a.Rn || !a.scale || !a.X || !ints)
        return fail(PMLG_ERR_ALLOC, "pmbasis: device arena exhausted");
    a.piv = reinterpret_cast<int *>(ints);
    a.rowinfo = a.piv + m;
    a.ctl = a.rowinfo + m;
    a.shift = dShift;
    a.mc = ctx->mc;
    const int64_t lt = lenF < ordL ? lenF : ordL;
    hipLaunchKernelGGL(leaf_init_kernel, dim3(grid_for((int64_t)m * m * cap + (int64_t)m * n * ordL)),
                       dim3(256), 0, s, a, F, strideF, lt);
    HIP_TRY(hipGetLastError(), "leaf_init_kernel");
    // p = 2 has no Montgomery form (p must be odd): plain residues as for p >= 2^62
    const bool big = ctx->p >= ((uint64_t)1 << 62) || ctx->p == 2;
    const bool small = !big && ctx->p < ((uint64_t)1 << 30);
    const size_t lds = step_lds_bytes(m, n, small ? 4 : 8);
    if (lds > kStepLdsMax)
        return fail(PMLG_ERR_UNSUPPORTED, "M-Basis: an m x n residual of this size does not fit "
                                          "the step kernel's LDS");
    static bool attr = false;
    if (!attr) {
        const int maxlds = (int)kStepLdsMax;
        const void *kernels[] = {(const void *)mbasis_step_kernel<Mont64Arith, 1>,
                                 (const void *)mbasis_step_kernel<Mont64Arith, 2>,
                                 (const void *)mbasis_step_kernel<Mont32Arith, 1>,
                                 (const void *)mbasis_step_kernel<Mont32Arith, 2>,
                                 (const void *)mbasis_step_kernel<BigArith, 1>,
                                 (const void *)mbasis_step_kernel<BigArith, 2>};
        for (const void *k : kernels)
            HIP_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, maxlds),
                    "hipFuncSetAttribute(mbasis_step_kernel)");
        attr = true;
    }
    const unsigned ugrid = grid_for((int64_t)m * m * cap + (int64_t)m * n * ordL);
    // rows of the transposed residual per lane of the composing wave
    const bool one_row = n <= 64;
    for (int ord = 0; ord < ordL; ++ord) {
        if (small) {
            if (one_row)
                hipLaunchKernelGGL((mbasis_step_kernel<Mont32Arith, 1>), dim3(1), dim3(kStepThreads), lds,
                                   s, a, ord);
            else
                hipLaunchKernelGGL((mbasis_step_kernel<Mont32Arith, 2>), dim3(1), dim3(kStepThreads), lds,
                                   s, a, ord);
            hipLaunchKernelGGL((mbasis_update_kernel<true, false>), dim3(ugrid), dim3(256), 0, s, a,
                               ord);
        } else if (big) {
            if (one_row)
                hipLaunchKernelGGL((mbasis_step_kernel<BigArith, 1>), dim3(1), dim3(kStepThreads), lds,
                                   s, a, ord);
            else
                hipLaunchKernelGGL((mbasis_step_kernel<BigArith, 2>), dim3(1), dim3(kStepThreads), lds,
                                   s, a, ord);
            hipLaunchKernelGGL((mbasis_update_kernel<false, true>), dim3(ugrid), dim3(256), 0, s, a,
                               ord);
        } else {
            if (one_row)
                hipLaunchKernelGGL((mbasis_step_kernel<Mont64Arith, 1>), dim3(1), dim3(kStepThreads), lds,
                                   s, a, ord);
            else
                hipLaunchKernelGGL((mbasis_step_kernel<Mont64Arith, 2>), dim3(1), dim3(kStepThreads), lds,
                                   s, a, ord);
            hipLaunchKernelGGL((mbasis_update_kernel<false, false>), dim3(ugrid), dim3(256), 0, s, a,
                               ord);
        }
    }
    HIP_TRY(hipGetLastError(), "mbasis kernels");
    int ctl[C_WORDS];
    HIP_TRY(hipMemcpyAsync(ctl, a.ctl, sizeof(ctl), hipMemcpyDeviceToHost, s), "D2H(ctl)");
    HIP_TRY(hipStreamSynchronize(s), "leaf sync");
    const int res = ctl[C_CUR] ^ ctl[C_FLIP];
    const int L = ctl[C_LEN] + ctl[C_GROW];
    if (ordL > 0) {
        // divide the row scales out of the basis
        if (small)
            hipLaunchKernelGGL(leaf_scale_inv_kernel<Mont32Arith>, dim3(1), dim3(kMaxM), 0, s, a);
        else if (big)
            hipLaunchKernelGGL(leaf_scale_inv_kernel<BigArith>, dim3(1), dim3(kMaxM), 0, s, a);
        else
            hipLaunchKernelGGL(leaf_scale_inv_kernel<Mont64Arith>, dim3(1), dim3(kMaxM), 0, s, a);
        hipLaunchKernelGGL(leaf_rescale_kernel, dim3(grid_for((int64_t)m * m * L)), dim3(256), 0, s,
                           a.app[res], (int64_t)cap, m, (int64_t)L, a.scale, ctx->mc);
        HIP_TRY(hipGetLastError(), "leaf_rescale_kernel");
    }
    // result to the mark (app[0] sits there; move app[1] down if needed)
    out->d = a.app[0];
    out->cap = cap;
    out->len = L;
    if (res == 1) {
        hipLaunchKernelGGL(copy_poly_kernel, dim3(grid_for((int64_t)m * m * L)), dim3(256), 0, s,
                           a.app[0], (int64_t)cap, a.app[1], (int64_t)cap, (int64_t)m * m, (int64_t)L);
        HIP_TRY(hipGetLastError(), "copy_poly_kernel");
    }
    ar.top = mark + (((size_t)m * m * cap * 8 + 255) & ~(size_t)255);
    return PMLG_OK;
}

// PM-Basis recursion (approximant_basis.c:33-69); result at the arena mark.
pmlg_status pmb_rec(pmlg_ctx ctx, Arena &ar, const uint64_t *F, int64_t strideF, int64_t lenF,
                    int m, int n, int64_t order, int64_t *dShift, int method, int thres, Poly *out,
                    hipStream_t s)
{
    const int64_t lt = lenF < order ? lenF : order;
    if (order <= thres)
        return leaf(ctx, ar, F, strideF, lt, m, n, (int)order, dShift, out, s);
    const size_t mark = ar.top;
    const int64_t order1 = order >> 1, order2 = order - order1;
    Poly P1;
    pmlg_status st = pmb_rec(ctx, ar, F, strideF, lt, m, n, order1, dShift, method, thres, &P1, s);
    if (st != PMLG_OK)
        return st;
    // residual = (P1 * F mod x^order) div x^order1
    uint64_t *Rz = ar.alloc((size_t)m * n * order2);
    if (!Rz)
        return fail(PMLG_ERR_ALLOC, "pmbasis: device arena exhausted");
    st = mulmid_dev_strided(ctx, Rz, order2, P1.d, P1.cap, F, strideF, m, m, n, P1.len, lt, order1,
                            order, method, s);
    if (st != PMLG_OK)
        return st;
    Poly P2;
    st = pmb_rec(ctx, ar, Rz, order2, order2, m, n, order2, dShift, method, thres, &P2, s);
    if (st != PMLG_OK)
        return st;
    // appbas = P2 * P1
    const int64_t L = P1.len + P2.len - 1;
    uint64_t *P = ar.alloc((size_t)m * m * L);
    if (!P)
        return fail(PMLG_ERR_ALLOC, "pmbasis: device arena exhausted");
    st = mulmid_dev_strided(ctx, P, L, P2.d, P2.cap, P1.d, P1.cap, m, m, m, P2.len, P1.len, 0, L,
                            method, s);
    if (st != PMLG_OK)
        return st;
    // move the result down to the mark (P1, Rz and P2 are dead)
    uint64_t *dst = reinterpret_cast<uint64_t *>(ar.base + mark);
    hipLaunchKernelGGL(copy_poly_kernel, dim3(grid_for((int64_t)m * m * L)), dim3(256), 0, s, dst, L,
                       P, L, (int64_t)m * m, L);
    HIP_TRY(hipGetLastError(), "copy_poly_kernel");
    ar.top = mark + (((size_t)m * m * L * 8 + 255) & ~(size_t)255);
    out->d = dst;
    out->cap = L;
    out->len = L;
    return PMLG_OK;
}

size_t arena_bytes(int64_t m, int64_t n, int64_t order, int64_t ordleaf)
{
    // generous bound: a few copies of every level's operands
    const double lvl = (double)m * m * (order + 2) * 4 + (double)m * n * (order + 1) * 2;
    const double leafb =
        (double)m * m * (ordleaf + 2) * 2 + (double)m * n * (ordleaf + 2) * 2 + m * m + 64;
    return (size_t)((2.5 * lvl + leafb + 4096.0) * 8.0);
}

// M-Basis / PM-Basis crossover.  The basis does not depend on it: the second
// half of the recursion repeats, on the residual of the first half's basis,
// the steps that M-Basis would take from that basis on (each step depends only
// on the current residual coefficient and shifts, and the basis updates are
// row operations that commute with the product by the first half's basis), so
// the leaf order is a pure speed choice.  Threshold t against 2t differs only
// at the nodes of order 2t, split into two leaves of order t plus a middle
// product and a product, or solved as one leaf: both are timed on seeded
// synthetic input of the caller's shape, first for t = 32 (the reference's
// threshold), then for t = 16 if the split won there, else for t = 64, and
// then for t = 128 if the leaf of 128 won too.  Measured once per context and
// (m, n, method), for orders of at least kTuneMinOrder (below it the
// reference's threshold is used as it is).
constexpr int kMaxLeafOrder = 256;
constexpr int64_t kTuneMinOrder = 512;

int64_t tune_leaf_bytes(int64_t m, int64_t n)
{
    return (int64_t)arena_bytes(m, n, kMaxLeafOrder, kMaxLeafOrder) + (m * n * kMaxLeafOrder + m) * 8 + 4096;
}

pmlg_status tune_leaf_order(pmlg_ctx ctx, int m, int n, int method, hipStream_t s, int *out)
{
    const uint64_t key = ((uint64_t)m << 16) | ((uint64_t)n << 8) | (uint64_t)(method & 0xff);
    auto it = ctx->leaf_orders.find(key);
    if (it != ctx->leaf_orders.end()) {
        *out = it->second;
        return PMLG_OK;
    }
    HIP_TRY(ctx->arena.ensure((size_t)tune_leaf_bytes(m, n)), "hipMalloc(arena)");
    Arena ar;
    ar.base = static_cast<char *>(ctx->arena.ptr);
    ar.size = ctx->arena.bytes;
    const int64_t len = kMaxLeafOrder;
    uint64_t *dF = ar.alloc((size_t)(m * n * len));
    int64_t *dShift = reinterpret_cast<int64_t *>(ar.alloc((size_t)m));
    if (!dF || !dShift)
        return fail(PMLG_ERR_ALLOC, "pmbasis: device arena exhausted");
    std::vector<uint64_t> F((size_t)(m * n * len));
    uint64_t x = 0x9e3779b97f4a7c15ull ^ key;
    for (uint64_t &v : F) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        v = x % ctx->p;
    }
    HIP_TRY(hipMemcpyAsync(dF, F.data(), F.size() * 8, hipMemcpyHostToDevice, s), "H2D(tuning input)");
    const std::vector<int64_t> zero((size_t)m, 0);
    const size_t mark = ar.top;
    // best of three runs (the first also loads the transform tables)
    auto run = [&](int64_t order, int thres, double *sec) -> pmlg_status {
        *sec = 1e30;
        for (int r = 0; r < 3; ++r) {
            HIP_TRY(hipMemcpyAsync(dShift, zero.data(), (size_t)m * 8, hipMemcpyHostToDevice, s),
                    "H2D(shift)");
            HIP_TRY(hipStreamSynchronize(s), "sync");
            const auto t0 = std::chrono::steady_clock::now();
            Poly P;
            pmlg_status st = pmb_rec(ctx, ar, dF, len, len, m, n, order, dShift, method, thres, &P, s);
            if (st != PMLG_OK)
                return st;
            HIP_TRY(hipStreamSynchronize(s), "sync");
            *sec = std::min(*sec, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
            ar.top = mark;
        }
        return PMLG_OK;
    };
    // true when two leaves of order t (and the products) beat one of order 2t
    auto split_wins = [&](int t, bool *wins) -> pmlg_status {
        double leaf_s = 0, split_s = 0;
        pmlg_status st = run(2 * t, 2 * t, &leaf_s);
        if (st == PMLG_OK)
            st = run(2 * t, t, &split_s);
        *wins = split_s < leaf_s;
        return st;
    };
    bool w32 = false, w = false;
    pmlg_status st = split_wins(32, &w32);
    if (st == PMLG_OK)
        st = split_wins(w32 ? 16 : 64, &w);
    if (st != PMLG_OK)
        return st;
    int best = w32 ? (w ? 16 : 32) : (w ? 64 : 0);
    if (best == 0) {
        st = split_wins(128, &w);
        if (st != PMLG_OK)
            return st;
        best = w ? 128 : 256;
    }
    ctx->leaf_orders[key] = best;
    *out = best;
    return PMLG_OK;
}

}  // namespace

pmlg_status basis_common(pmlg_ctx ctx, uint64_t *dP, int64_t capP, int64_t *lenP,
                         int64_t *shift, const uint64_t *dF, int64_t strideF, int64_t m, int64_t n,
                         int64_t lenF, int64_t order, int method, bool iterative, hipStream_t s,
                         uint64_t *hP)
{
    if (m <= 0 || n <= 0 || m > kMaxM || n > kMaxN || order < 0 || lenF < 0)
        return fail(PMLG_ERR_ARG, "approximant basis: 1 <= m <= 256, 1 <= n <= 128, order >= 0");
    if (step_lds_bytes((int)m, (int)n, ctx->p < ((uint64_t)1 << 30) && ctx->p != 2 ? 4 : 8) >
        kStepLdsMax)
        return fail(PMLG_ERR_UNSUPPORTED,
                    "approximant basis: the m x n residual exceeds one workgroup's LDS (n (m | 1) "
                    "words plus the step's buffers: e.g. m = 256 allows n <= 62 for p >= 2^30)");
    if (!ctx->is_prime && ctx->p != 2)
        return fail(PMLG_ERR_UNSUPPORTED, "approximant bases need a prime modulus");
    if (capP < order + 1)
        return fail(PMLG_ERR_ARG, "approximant basis: output capacity must be at least order + 1");
    int thres = ctx->leaf_order > 0 ? ctx->leaf_order : kPmbasisThres;
    if (!iterative && ctx->leaf_order == 0 && order >= kTuneMinOrder) {
        pmlg_status st = tune_leaf_order(ctx, (int)m, (int)n, method, s, &thres);
        if (st != PMLG_OK)
            return st;
    }
    // the M-Basis kernels index one leaf's items in 32 bits
    const int64_t ordleaf = iterative ? order : std::min<int64_t>(order, thres);
    if (m * m * (ordleaf + 1) + m * n * ordleaf >= ((int64_t)1 << 31))
        return fail(PMLG_ERR_UNSUPPORTED, "M-Basis: m^2 (order + 1) + m n order must stay below 2^31 "
                                          "(use PM-Basis for this order)");
    // device arena
    size_t need = arena_bytes(m, n, order, ordleaf) +
                  (size_t)(m * m * (ordleaf + 2) * 2 + m * n * (ordleaf + 1) * 2) * 8;
    HIP_TRY(ctx->arena.ensure(need), "hipMalloc(arena)");
    Arena ar;
    ar.base = static_cast<char *>(ctx->arena.ptr);
    ar.size = ctx->arena.bytes;
    int64_t *dShift = reinterpret_cast<int64_t *>(ar.alloc((size_t)m));
    HIP_TRY(hipMemcpyAsync(dShift, shift, (size_t)m * 8, hipMemcpyHostToDevice, s), "H2D(shift)");
    Poly P;
    pmlg_status st;
    if (iterative)
        st = leaf(ctx, ar, dF, strideF, std::min(lenF, order), (int)m, (int)n, (int)order, dShift, &P,
                  s);
    else
        st = pmb_rec(ctx, ar, dF, strideF, lenF, (int)m, (int)n, order, dShift, method, thres, &P, s);
    if (st != PMLG_OK)
        return st;
    HIP_TRY(hipMemcpyAsync(shift, dShift, (size_t)m * 8, hipMemcpyDeviceToHost, s), "D2H(shift)");
    const int64_t L = std::min<int64_t>(P.len, capP);
    if (hP) {
        // entries packed to the computed length on the device (stride L)
        const uint64_t *src = P.cap == L ? P.d : nullptr;
        if (!src) {
            uint64_t *packed = ar.alloc((size_t)(m * m * L));
            if (packed) {
                hipLaunchKernelGGL(copy_poly_kernel, dim3(grid_for(m * m * L)), dim3(256), 0, s, packed,
                                   (int64_t)L, P.d, P.cap, (int64_t)(m * m), L);
                HIP_TRY(hipGetLastError(), "copy_poly_kernel");
                src = packed;
            }
        }
        int64_t Ln = 0;
        if (src) {
            // through the pinned staging chunks: each entry trimmed of its zero
            // top coefficients and zero-filled up to capP; the normalised
            // length is the longest entry's (the basis has degree at most order)
            std::vector<uint64_t *> ptr((size_t)(m * m));
            std::vector<int64_t> len((size_t)(m * m));
            for (int64_t e = 0; e < m * m; ++e)
                ptr[(size_t)e] = hP + e * capP;
            pmlg_status st2 = stage_download(ctx, ptr.data(), len.data(), src, m * m, L, s, capP);
            if (st2 != PMLG_OK)
                return st2;
            HIP_TRY(hipStreamSynchronize(s), "sync");  // the shifts' read-back
            for (int64_t e = 0; e < m * m; ++e)
                Ln = std::max(Ln, len[(size_t)e]);
        } else {
            // no arena room to pack: one transfer of the entries at capacity
            std::vector<uint64_t> tmp((size_t)(m * m * P.cap));
            HIP_TRY(hipMemcpyAsync(tmp.data(), P.d, tmp.size() * 8, hipMemcpyDeviceToHost, s), "D2H(P)");
            HIP_TRY(hipStreamSynchronize(s), "sync");
            for (int64_t e = 0; e < m * m; ++e)
                for (int64_t d = L - 1; d >= Ln; --d)
                    if (tmp[(size_t)(e * P.cap + d)]) {
                        Ln = d + 1;
                        break;
                    }
            for (int64_t e = 0; e < m * m; ++e) {
                std::memcpy(hP + e * capP, tmp.data() + e * P.cap, (size_t)Ln * 8);
                std::memset(hP + e * capP + Ln, 0, (size_t)(capP - Ln) * 8);
            }
        }
        *lenP = Ln;
    } else {
        hipLaunchKernelGGL(copy_poly_kernel, dim3(grid_for(m * m * L)), dim3(256), 0, s, dP, capP, P.d,
                           P.cap, (int64_t)(m * m), L);
        HIP_TRY(hipGetLastError(), "copy_poly_kernel");
        HIP_TRY(hipStreamSynchronize(s), "sync");
        *lenP = L;
    }
    return PMLG_OK;
}

}  // namespace pmlg

using namespace pmlg;

extern "C" {

pmlg_status pmlg_pmbasis_dev(pmlg_ctx ctx, uint64_t *dP, int64_t capP, int64_t *lenP,
                             int64_t *shift, const uint64_t *dF, int64_t m, int64_t n,
                             int64_t lenF, int64_t order, int method, void *stream)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    return basis_common(ctx, dP, capP, lenP, shift, dF, lenF, m, n, lenF, order, method, false, s,
                        nullptr);
}

static pmlg_status basis_host(pmlg_ctx ctx, uint64_t *P, int64_t *lenP, int64_t *shift,
                              const uint64_t *F, int64_t m, int64_t n, int64_t lenF, int64_t order,
                              int64_t capP, int method, bool iterative)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    if (m <= 0 || n <= 0 || lenF < 0 || !P || !lenP || !shift || (lenF > 0 && !F))
        return fail(PMLG_ERR_ARG, "approximant basis: bad arguments");
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    hipStream_t s = ctx->stream;
    const size_t fb = (size_t)(m * n * (lenF > 0 ? lenF : 1)) * 8;
    HIP_TRY(ctx->in_a.ensure(fb), "hipMalloc(F)");
    if (lenF > 0) {
        // through the pinned staging chunks, entry by entry
        std::vector<const uint64_t *> ptr((size_t)(m * n));
        std::vector<int64_t> len((size_t)(m * n), lenF);
        for (int64_t e = 0; e < m * n; ++e)
            ptr[(size_t)e] = F + e * lenF;
        st = stage_upload(ctx, static_cast<uint64_t *>(ctx->in_a.ptr), ptr.data(), len.data(), m * n,
                          lenF);
        if (st != PMLG_OK)
            return st;
    }
    return basis_common(ctx, nullptr, capP, lenP, shift, static_cast<const uint64_t *>(ctx->in_a.ptr),
                        lenF, m, n, lenF, order, method, iterative, s, P);
}

pmlg_status pmlg_pmbasis(pmlg_ctx ctx, uint64_t *P, int64_t *lenP, int64_t *shift,
                         const uint64_t *F, int64_t m, int64_t n, int64_t lenF, int64_t order,
                         int64_t capP, int method)
{
    return basis_host(ctx, P, lenP, shift, F, m, n, lenF, order, capP, method, false);
}

pmlg_status pmlg_mbasis(pmlg_ctx ctx, uint64_t *P, int64_t *lenP, int64_t *shift,
                        const uint64_t *F, int64_t m, int64_t n, int64_t lenF, int64_t order,
                        int64_t capP)
{
    return basis_host(ctx, P, lenP, shift, F, m, n, lenF, order, capP, PMLG_METHOD_AUTO, true);
}

pmlg_status pmlg_pmbasis_set_leaf_order(pmlg_ctx ctx, int64_t order)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    if (order < 0 || order > kMaxLeafOrder)
        return fail(PMLG_ERR_ARG, "pmbasis leaf order must be 0 (measured) or in [1, 256]");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->leaf_order = (int)order;
    return PMLG_OK;
}

pmlg_status pmlg_pmbasis_leaf_order(pmlg_ctx ctx, int64_t m, int64_t n, int method, int64_t *order)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    if (m <= 0 || n <= 0 || m > kMaxM || n > kMaxN || !order)
        return fail(PMLG_ERR_ARG, "pmbasis leaf order: 1 <= m <= 256, 1 <= n <= 128");
    if (!ctx->is_prime && ctx->p != 2)
        return fail(PMLG_ERR_UNSUPPORTED, "approximant bases need a prime modulus");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->leaf_order > 0) {
        *order = ctx->leaf_order;
        return PMLG_OK;
    }
    DeviceGuard g(ctx->device);
    int t = kPmbasisThres;
    st = tune_leaf_order(ctx, (int)m, (int)n, method, ctx->stream, &t);
    *order = t;
    return st;
}

}  // extern "C"

// ===========================================================================
// Constant matrices and evaluation (used by the approximant-basis certificate)
// ===========================================================================
namespace pmlg {
namespace {

// Gaussian elimination of an r x c matrix held in device memory, by one
// workgroup (the matrix is modified in place; fcol holds r words of scratch):
// rank, and the determinant when r == c (product of the pivots, sign of the
// row swaps).  Barriers order the global-memory updates of one workgroup.
__global__ void __launch_bounds__(1024) const_gauss_kernel(uint64_t *A, uint64_t *fcol, int r, int c,
                                                           int *out_rank, uint64_t *out_det,
                                                           ModConst mc)
{
    __shared__ int s_pr;
    __shared__ uint64_t s_inv;
    const int tid = threadIdx.x;
    const uint64_t p = mc.p;
    int rank = 0;
    uint64_t det = 1 % p;
    bool neg = false;
    for (int col = 0; col < c && rank < r; ++col) {
        if (tid == 0) {
            int pr = -1;
            for (int i = rank; i < r; ++i)
                if (A[(size_t)i * c + col]) {
                    pr = i;
                    break;
                }
            s_pr = pr;
            if (pr >= 0)
                s_inv = invmod_g(A[(size_t)pr * c + col], mc);
        }
        __syncthreads();
        const int pr = s_pr;
        if (pr < 0) {
            det = 0;
            __syncthreads();
            continue;
        }
        if (pr != rank) {
            for (int j = tid; j < c; j += blockDim.x) {
                const uint64_t t = A[(size_t)pr * c + j];
                A[(size_t)pr * c + j] = A[(size_t)rank * c + j];
                A[(size_t)rank * c + j] = t;
            }
            neg = !neg;
        }
        __syncthreads();
        det = mulmod_g(det, A[(size_t)rank * c + col], mc);
        for (int i = tid; i < r; i += blockDim.x)
            fcol[i] = mulmod_g(A[(size_t)i * c + col], s_inv, mc);
        __syncthreads();
        for (int64_t idx = (int64_t)(rank + 1) * c + tid; idx < (int64_t)r * c; idx += blockDim.x) {
            const int i = (int)(idx / c), j = (int)(idx - (int64_t)i * c);
            A[idx] = submod_g(A[idx], mulmod_g(fcol[i], A[(size_t)rank * c + j], mc), p);
        }
        ++rank;
        __syncthreads();
    }
    if (tid == 0) {
        *out_rank = rank;
        if (out_det)
            *out_det = (r == c && rank == r) ? (neg && det ? p - det : det) : 0;
    }
}

// out[e] = sum_d A[e*len + d] x^d (Horner), one thread per entry
__global__ void eval_kernel(uint64_t *out, const uint64_t *A, int64_t entries, int64_t len,
                            uint64_t x, ModConst mc)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < entries;
         e += (int64_t)gridDim.x * blockDim.x) {
        uint64_t v = 0;
        for (int64_t d = len - 1; d >= 0; --d) {
            v = addmod_any(mulmod_g(v, x, mc), A[e * len + d], mc.p);
        }
        out[e] = v;
    }
}

}  // namespace

pmlg_status const_rank_dev(pmlg_ctx ctx, uint64_t *dM, uint64_t *dScratch, int64_t r, int64_t c,
                           int64_t *rank, hipStream_t s)
{
    int *drank = reinterpret_cast<int *>(dScratch + r);
    hipLaunchKernelGGL(const_gauss_kernel, dim3(1), dim3(1024), 0, s, dM, dScratch, (int)r, (int)c,
                       drank, nullptr, ctx->mc);
    HIP_TRY(hipGetLastError(), "const_gauss_kernel");
    int hr = 0;
    HIP_TRY(hipMemcpyAsync(&hr, drank, sizeof(int), hipMemcpyDeviceToHost, s), "D2H(rank)");
    HIP_TRY(hipStreamSynchronize(s), "sync");
    *rank = hr;
    return PMLG_OK;
}

}  // namespace pmlg

extern "C" {

pmlg_status pmlg_const_rank_det(pmlg_ctx ctx, const uint64_t *M, int64_t r, int64_t c,
                                int64_t *rank, uint64_t *det)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    if (r < 0 || c < 0 || r > (1 << 16) || c > (1 << 16) || r * c > (1 << 22) || !rank ||
        (r * c > 0 && !M))
        return fail(PMLG_ERR_ARG, "const_rank_det: at most 2^22 entries");
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    hipStream_t s = ctx->stream;
    const size_t words = (size_t)(r * c) + (size_t)r + 4;
    HIP_TRY(ctx->in_b.ensure(words * 8), "hipMalloc");
    uint64_t *dM = static_cast<uint64_t *>(ctx->in_b.ptr);
    uint64_t *dcol = dM + r * c;
    uint64_t *ddet = dcol + r;
    int *drank = reinterpret_cast<int *>(ddet + 1);
    if (r * c)
        HIP_TRY(hipMemcpyAsync(dM, M, (size_t)(r * c) * 8, hipMemcpyHostToDevice, s), "H2D");
    hipLaunchKernelGGL(const_gauss_kernel, dim3(1), dim3(1024), 0, s, dM, dcol, (int)r, (int)c,
                       drank, ddet, ctx->mc);
    HIP_TRY(hipGetLastError(), "const_gauss_kernel");
    int hr = 0;
    uint64_t hd = 0;
    HIP_TRY(hipMemcpyAsync(&hr, drank, sizeof(int), hipMemcpyDeviceToHost, s), "D2H");
    HIP_TRY(hipMemcpyAsync(&hd, ddet, 8, hipMemcpyDeviceToHost, s), "D2H");
    HIP_TRY(hipStreamSynchronize(s), "sync");
    *rank = hr;
    if (det)
        *det = hd;
    return PMLG_OK;
}

pmlg_status pmlg_polmat_eval(pmlg_ctx ctx, uint64_t *out, const uint64_t *A, int64_t r, int64_t c,
                             int64_t len, uint64_t x)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    if (r < 0 || c < 0 || len < 0 || !out || (r * c * len > 0 && !A))
        return fail(PMLG_ERR_ARG, "polmat_eval: bad arguments");
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    hipStream_t s = ctx->stream;
    const size_t ab = (size_t)(r * c * len) * 8, ob = (size_t)(r * c) * 8;
    HIP_TRY(ctx->in_a.ensure(ab ? ab : 8), "hipMalloc");
    HIP_TRY(ctx->out_c.ensure(ob ? ob : 8), "hipMalloc");
    if (ab)
        HIP_TRY(hipMemcpyAsync(ctx->in_a.ptr, A, ab, hipMemcpyHostToDevice, s), "H2D");
    hipLaunchKernelGGL(eval_kernel, dim3(grid_for(r * c)), dim3(256), 0, s,
                       static_cast<uint64_t *>(ctx->out_c.ptr),
                       static_cast<const uint64_t *>(ctx->in_a.ptr), r * c, len, x % ctx->p, ctx->mc);
    HIP_TRY(hipGetLastError(), "eval_kernel");
    if (ob)
        HIP_TRY(hipMemcpyAsync(out, ctx->out_c.ptr, ob, hipMemcpyDeviceToHost, s), "D2H");
    HIP_TRY(hipStreamSynchronize(s), "sync");
    return PMLG_OK;
}

}  // extern "C"
