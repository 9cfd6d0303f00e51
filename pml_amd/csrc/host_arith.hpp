// Host-side number theory for context set-up: primality, roots of unity,
// Shoup/Barrett constants.  Exact 128-bit arithmetic.
#pragma once

#include <stdint.h>

#include "modarith.hpp"

namespace pmlg {

typedef unsigned __int128 u128_t;

static inline uint64_t host_mulmod(uint64_t a, uint64_t b, uint64_t p)
{
    return (uint64_t)(((u128_t)a * b) % p);
}

static inline uint64_t host_powmod(uint64_t a, uint64_t e, uint64_t p)
{
    uint64_t r = 1 % p;
    a %= p;
    while (e) {
        if (e & 1)
            r = host_mulmod(r, a, p);
        a = host_mulmod(a, a, p);
        e >>= 1;
    }
    return r;
}

static inline uint64_t host_invmod(uint64_t a, uint64_t p)  // p prime
{
    return host_powmod(a, p - 2, p);
}

// Miller-Rabin with a witness set that is exact for all n < 2^64.
static inline bool host_is_prime(uint64_t n)
{
    if (n < 2)
        return false;
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t q : small) {
        if (n == q)
            return true;
        if (n % q == 0)
            return false;
    }
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) {
        d >>= 1;
        ++s;
    }
    static const uint64_t wit[] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    for (uint64_t a0 : wit) {
        uint64_t a = a0 % n;
        if (!a)
            continue;
        uint64_t x = host_powmod(a, d, n);
        if (x == 1 || x == n - 1)
            continue;
        bool comp = true;
        for (int r = 1; r < s; ++r) {
            x = host_mulmod(x, x, n);
            if (x == n - 1) {
                comp = false;
                break;
            }
        }
        if (comp)
            return false;
    }
    return true;
}

// A primitive 2^k-th root of unity mod p (k <= v2(p-1)), from the first
// quadratic non-residue g: g^((p-1)/2^k) has order exactly 2^k.
static inline uint64_t host_root_of_unity_2k(uint64_t p, int k)
{
    if (k == 0)
        return 1;
    for (uint64_t g = 2; g < p; ++g) {
        if (host_powmod(g, (p - 1) / 2, p) == p - 1)
            return host_powmod(g, (p - 1) >> k, p);
    }
    return 0;
}

static inline uint64_t host_shoup_pre64(uint64_t w, uint64_t p)
{
    return (uint64_t)(((u128_t)w << 64) / p);
}

static inline uint64_t host_shoup_pre32(uint64_t w, uint64_t p)  // p < 2^32
{
    return (uint64_t)((w << 32) / p);
}

static inline uint64_t host_pow2_mod(int e, uint64_t p)
{
    return host_powmod(2, (uint64_t)e, p);
}

static inline ModConst host_modconst(uint64_t p)
{
    ModConst mc{};
    mc.p = p;
    mc.p2 = 2 * p;
    mc.bm = (uint64_t)(((u128_t)1 << 64) / p);
    mc.c64 = (uint64_t)(((u128_t)1 << 64) % p);
    mc.c64q = host_shoup_pre64(mc.c64, p);
    mc.c40 = host_pow2_mod(40, p);
    mc.c40q = host_shoup_pre64(mc.c40, p);
    mc.c80 = host_pow2_mod(80, p);
    mc.c80q = host_shoup_pre64(mc.c80, p);
    mc.c32 = host_pow2_mod(32, p);
    mc.c32q = host_shoup_pre64(mc.c32, p);
    {
        // -p^-1 mod 2^64 by Newton iteration (p odd)
        uint64_t inv = p;  // correct to 3 bits
        for (int i = 0; i < 6; ++i)
            inv *= 2 - p * inv;
        mc.pinvneg = (uint64_t)0 - inv;
    }
    mc.r64 = mc.c64;
    mc.r64q = mc.c64q;
    mc.redc_small = 0;
    const uint64_t two61 = (uint64_t)1 << 61;
    mc.goff = p * ((two61 + p - 1) / p);
    int bits = 0;
    while (bits < 64 && (p >> bits))
        ++bits;
    mc.bits = bits;
    if (p < ((uint64_t)1 << 30))
        mc.limbs = 4;
    else if (p < ((uint64_t)1 << 61))
        mc.limbs = 8;
    else
        mc.limbs = 0;  // MFMA path unavailable
    mc.p32 = (uint32_t)p;
    mc.p32x2 = (uint32_t)(2 * p);
    mc.norm = __builtin_clzll(p);
    mc.dnorm = p << mc.norm;
    mc.dinv = (uint64_t)((~(u128_t)0) / mc.dnorm - ((u128_t)1 << 64));
    return mc;
}

}  // namespace pmlg
