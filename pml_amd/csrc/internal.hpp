// Internal declarations shared by the device translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "modarith.hpp"

namespace pmlg {

// Root-of-unity tables for one (p, log2 N): tw[t] = w^t, itw[t] = w^-t for
// t < N, with their Shoup companions; w a primitive N-th root of unity.
template <typename W>
struct NttTab {
    const W *tw, *twq, *itw, *itwq;
    W ninv, ninvq;  // N^-1 mod p
    // Per-stage tables of the LDS sub-transform (size M = 2^logM, root
    // w_M = w^(N/M)): stage s holds (w_M^(x 2^s), Shoup companion) pairs for
    // x < 2^(logM-1-s), stored interleaved from offset 2*(M - 2^(logM-s)).
    const W *stw, *sitw;
    W ninvr, ninvrq;  // N^-1 * 2^64 mod p: scaling after a Montgomery (REDC) pointwise product
};

// Largest transform held in LDS by one workgroup (elements).
template <typename W>
struct NttLimits;
template <>
struct NttLimits<uint64_t> {
    static constexpr int max_log_m = 13;  // 64 KiB of LDS
};
template <>
struct NttLimits<uint32_t> {
    static constexpr int max_log_m = 14;  // 64 KiB of LDS
};

// Tuning knobs of the transforms (picked per context, see the autotuner).
struct NttTune {
    int fwd_log_e = 5;  // log2 elements per thread in the register kernels (4 or 5)
    int inv_log_e = 4;
};

// Evaluation arrays use a point-blocked layout: position pos of entry e (of
// `count` entries) is at ((pos >> logP) * count + e) * 2^logP + (pos mod 2^logP).

// Forward transforms of `count` entries: entry e reads in[e*in_stride + i]
// for i < in_len (zero beyond), writes its N evaluations in blocked layout.
template <typename Win, typename W>
hipError_t launch_ntt_fwd(const Win *in, int64_t in_stride, int64_t in_len, W *out,
                          int64_t count, int logN, int logP, const NttTune &tune,
                          const NttTab<W> &tab, W p, hipStream_t s);

// Inverse transforms (blocked input): coefficients [lo, hi) of each entry
// (clipped to N), scaled by 1/N, written to out[e*out_stride + (i - lo)] as
// Wout in [0,p).  tmp must hold count*N words of W when N exceeds the LDS limit.
template <typename W, typename Wout>
hipError_t launch_ntt_inv(const W *in, Wout *out, int64_t out_stride, int64_t lo, int64_t hi,
                          int64_t count, int logN, int logP, const NttTune &tune,
                          const NttTab<W> &tab, W p, W *tmp, hipStream_t s);

enum PointwiseKind { PW_PLAIN = 0, PW_MFMA = 1 };

// C_hat[i][j][q] = sum_k A_hat[i][k][q] * B_hat[k][j][q] mod p, q < N.
template <typename W>
hipError_t launch_pointwise(PointwiseKind kind, const W *Ah, const W *Bh, W *Ch, int m, int k,
                            int n, int64_t N, int logP, const ModConst &mc, hipStream_t s);

// Whether the int8-limb MFMA kernel supports this shape.
bool mfma_supported(int m, int k, int n, int limbs, uint64_t p);

}  // namespace pmlg
