// Modular arithmetic for word-size primes, shared by host and device code.
//
// Two word types are used on the device:
//   uint64_t for primes p < 2^62 (values kept lazily in [0, 2p) inside the NTT),
//   uint32_t for primes p < 2^30 (values kept lazily in [0, 2p)).
// Multiplications by a fixed operand (twiddles, recombination constants) use
// Shoup's precomputed quotient: for w < p and w' = floor(w * 2^B / p),
// q = hi_B(x * w'), r = x*w - q*p lies in [0, 2p) for every x < 2^B.
// General reductions use Barrett with m = floor(2^64 / p): for x < 2^64,
// q = hi64(x*m) and x - q*p lies in [0, 2p).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PMLG_HD __host__ __device__ __forceinline__
#else
#define PMLG_HD static inline
#endif

namespace pmlg {

PMLG_HD uint64_t mulhi64(uint64_t a, uint64_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

PMLG_HD uint32_t mulhi32(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

// ---------------------------------------------------------------------------
// Shoup products by a precomputed operand
// ---------------------------------------------------------------------------

// x * w mod p in [0, 2p), x < 2^64, w < p < 2^63, wq = floor(w 2^64 / p)
PMLG_HD uint64_t shoup_lazy(uint64_t x, uint64_t w, uint64_t wq, uint64_t p)
{
    uint64_t q = mulhi64(x, wq);
    return x * w - q * p;
}

PMLG_HD uint32_t shoup_lazy(uint32_t x, uint32_t w, uint32_t wq, uint32_t p)
{
    uint32_t q = mulhi32(x, wq);
    return x * w - q * p;
}

template <typename W>
PMLG_HD W reduce_2p(W x, W p)  // [0, 2p) -> [0, p)
{
    return x >= p ? x - p : x;
}

template <typename W>
PMLG_HD W shoup(W x, W w, W wq, W p)
{
    return reduce_2p<W>(shoup_lazy(x, w, wq, p), p);
}

// Lazy butterfly helpers on values in [0, 2p) with 4p < 2^bits(W).
template <typename W>
PMLG_HD W add_2p(W a, W b, W p2)  // (a + b) mod-ish, result in [0, 2p)
{
    W s = a + b;
    return s >= p2 ? s - p2 : s;
}

template <typename W>
PMLG_HD W sub_2p(W a, W b, W p2)  // a - b + 2p in [0, 4p)
{
    return a + p2 - b;
}

// ---------------------------------------------------------------------------
// Barrett reduction of a 64-bit value; m = floor(2^64 / p) (p >= 2)
// ---------------------------------------------------------------------------
PMLG_HD uint64_t barrett_lazy(uint64_t x, uint64_t m, uint64_t p)  // [0, 2p)
{
    uint64_t q = mulhi64(x, m);
    return x - q * p;
}

PMLG_HD uint64_t barrett(uint64_t x, uint64_t m, uint64_t p)  // [0, p)
{
    uint64_t r = barrett_lazy(x, m, p);
    return r >= p ? r - p : r;
}

// Per-prime constants handed to kernels by value.
struct ModConst {
    uint64_t p;
    uint64_t p2;        // 2p
    uint64_t bm;        // floor(2^64 / p)             (Barrett)
    uint64_t c64, c64q; // 2^64 mod p, Shoup companion
    // int8-limb recombination (see pointwise kernels)
    uint64_t c40, c40q; // 2^40 mod p
    uint64_t c80, c80q; // 2^80 mod p
    uint64_t c32, c32q; // 2^32 mod p
    uint64_t goff;      // multiple of p, >= 2^61 (offset making signed partial sums nonnegative)
    uint64_t pinvneg;   // -p^-1 mod 2^64 (Montgomery REDC)
    uint64_t r64, r64q; // 2^64 mod p = Montgomery factor R, Shoup companion
    int redc_small;     // set per call: REDC output already below 4p
    uint32_t p32, p32x2, p32q_unused;
    int bits;           // bit length of p
    int limbs;          // 8 for p < 2^62, 4 for p < 2^30
    // any p < 2^64: division of 128-bit values by the normalised modulus
    int norm;           // leading zero bits of p
    uint64_t dnorm;     // p << norm
    uint64_t dinv;      // floor((2^128 - 1) / dnorm) - 2^64
};

// Reduce a 128-bit value (hi, lo), hi < 2^64, to [0, p).
PMLG_HD uint64_t reduce128(uint64_t hi, uint64_t lo, const ModConst &mc)
{
    uint64_t h = barrett_lazy(hi, mc.bm, mc.p);            // [0, 2p)
    uint64_t t = shoup_lazy(h, mc.c64, mc.c64q, mc.p);     // [0, 2p)
    uint64_t l = barrett_lazy(lo, mc.bm, mc.p);            // [0, 2p)
    uint64_t s = t + l;                                    // [0, 4p) < 2^64
    s = s >= mc.p2 ? s - mc.p2 : s;
    return s >= mc.p ? s - mc.p : s;
}

// 64x64 -> 128 product (hi, lo).  On the device: four 32x32 -> 64
// multiply-adds whose partial sums give both halves, a b = lo32(ll)
// + 2^32 lo32(m2) + 2^64 (ah bh + hi32(m1) + hi32(m2)) with ll = al bl,
// m1 = al bh + hi32(ll), m2 = ah bl + lo32(m1) (no partial sum overflows);
// the separate high and low products take six.
PMLG_HD void mul128(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t al = (uint32_t)a, ah = (uint32_t)(a >> 32), bl = (uint32_t)b, bh = (uint32_t)(b >> 32);
    const uint64_t ll = (uint64_t)al * bl;
    const uint64_t m1 = (uint64_t)al * bh + (ll >> 32);
    const uint64_t m2 = (uint64_t)ah * bl + (uint32_t)m1;
    hi = (uint64_t)ah * bh + (m1 >> 32) + (m2 >> 32);
    lo = ((uint64_t)(uint32_t)m2 << 32) | (uint32_t)ll;
#else
    const unsigned __int128 r = (unsigned __int128)a * b;
    hi = (uint64_t)(r >> 64);
    lo = (uint64_t)r;
#endif
}

// (hi 2^64 + lo) mod p for any p < 2^64 and hi < p: the value and the
// modulus shifted left by `norm` (the top bit of the divisor set), then one
// division step with the precomputed reciprocal of Moller and Granlund
// ("Improved division by invariant integers", IEEE Trans. Computers 60(2),
// 2011, Algorithm 4), whose remainder is shifted back.
PMLG_HD uint64_t mod128_any(uint64_t hi, uint64_t lo, const ModConst &mc)
{
    const uint64_t u1 = mc.norm ? (hi << mc.norm) | (lo >> (64 - mc.norm)) : hi;
    const uint64_t u0 = lo << mc.norm;
    const uint64_t d = mc.dnorm;
    uint64_t q1, q0;
    mul128(mc.dinv, u1, q1, q0);
    q0 += u0;
    q1 += u1 + (q0 < u0 ? 1 : 0);
    q1 += 1;
    uint64_t r = u0 - q1 * d;
    if (r > q0)
        r += d;
    if (r >= d)
        r -= d;
    return r >> mc.norm;
}

// a * b mod p for residues a, b < p, any p < 2^64
PMLG_HD uint64_t mulmod_any(uint64_t a, uint64_t b, const ModConst &mc)
{
    uint64_t hi, lo;
    mul128(a, b, hi, lo);
    return mod128_any(hi, lo, mc);
}

// a + b mod p for residues a, b < p, any p < 2^64 (the sum may carry out)
PMLG_HD uint64_t addmod_any(uint64_t a, uint64_t b, uint64_t p)
{
    const uint64_t s = a + b;
    return (s < a || s >= p) ? s - p : s;
}

// 64x64 -> 128 multiply-accumulate into (hi, lo)
PMLG_HD void mac128(uint64_t &hi, uint64_t &lo, uint64_t a, uint64_t b)
{
    uint64_t h, l;
    mul128(a, b, h, l);
    lo += l;
    hi += h + (lo < l ? 1 : 0);
}

}  // namespace pmlg
