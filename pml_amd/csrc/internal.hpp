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
    // 3/4-truncated inverses: 2 N^-1 (2 N^-1 2^64 after REDC), w^(N/4) / 2
    // and w^(N/4)
    W ninv2, ninv2q, ninvr2, ninvr2q, thalf, thalfq, tquart, tquartq;
};

// Tables of the long transforms (N = R M with R > 4): the column pass runs
// R-point transforms over the M columns of each entry, then multiplies
// column i1, output g by w^(i1 g).  rt[k] = w_R^k and rti[k] = w_R^-k for
// k < R/2 (w_R = w^M); w^a = lo[a mod 2^slog] * hi[a >> slog] (and the same
// for w^-a with loi, hii), so the twist needs 2 sqrt(N) words, not N.
template <typename W>
struct ColTab {
    const W *rt, *rtq, *rti, *rtiq;
    const W *lo, *loq, *hi, *hiq;
    const W *loi, *loiq, *hii, *hiiq;
    int slog;
};

// Largest transform held in LDS by one workgroup (elements), and the
// largest transform at all: the column pass holds R x 4 words of one entry
// in 64 KiB of LDS.
template <typename W>
struct NttLimits;
template <>
struct NttLimits<uint64_t> {
    static constexpr int max_log_m = 13;  // 64 KiB of LDS
    static constexpr int max_log_r = 11;  // 2^11 x 4 columns x 8 bytes = 64 KiB
    static constexpr int max_log_n = max_log_m + max_log_r;
};
template <>
struct NttLimits<uint32_t> {
    static constexpr int max_log_m = 14;  // 64 KiB of LDS
    static constexpr int max_log_r = 12;  // 2^12 x 4 columns x 4 bytes = 64 KiB
    static constexpr int max_log_n = max_log_m + max_log_r;
};

// Tuning knobs of the transforms (picked per context, see the autotuner).
struct NttTune {
    int fwd_log_e = 5;  // log2 elements per thread in the register kernels (4 or 5)
    int inv_log_e = 4;
};

// Evaluation arrays use a point-blocked layout: position pos of entry e (of
// `count` entries) is at ((pos >> logP) * count + e) * 2^logP + (pos mod 2^logP).
// Both the transforms' evaluations (A_hat, B_hat) and the pointwise products
// (C_hat) use 32-byte runs: a 128-byte line holds four neighbouring entries
// at four points, which the pointwise product reads as one quad of entries
// and writes from one workgroup.  Measured on 16 products 32x32 * 32x32 of
// length 4096 (ms per launch, forward / pointwise / inverse):
//   32-byte runs for all                       1.34 / 1.00 / 0.82
//   128-byte runs for C_hat                    1.36 / 1.28 / 0.76
//   128-byte runs for A_hat, B_hat             1.34 / 1.08 / 0.80
// Whole-line runs let the inverse read (or the forward write) whole lines of
// one entry, but the pointwise product's 8-byte stores, or its quads of
// entries, then touch four times as many lines.
template <typename W>
inline int eval_log_run_in(int logN)
{
    const int want = sizeof(W) == 8 ? 2 : 3;
    return logN < want ? logN : want;
}

template <typename W>
inline int eval_log_run_out(int logN)
{
    return eval_log_run_in<W>(logN);
}

// 3/4-truncated transforms: a product of length <= 3N/4 needs only the
// evaluations at the first 3N/4 positions (the frequencies 0, 1 and 2 mod 4:
// the DIF output order of one sub-transform, or the first three of four
// sub-transforms when N exceeds one, split in four), for kTruncMinLog <=
// logN <= max_log_m + 2 (3N/4 then holds whole point blocks of 8).
constexpr int kTruncMinLog = 5;

template <typename W>
inline bool trunc_ok(int logN, int64_t lenC)
{
    return logN >= kTruncMinLog && logN <= NttLimits<W>::max_log_m + 2 &&
           lenC <= 3 * (((int64_t)1 << logN) >> 2);
}

// Forward transforms of `count` entries: entry e reads in[e*in_stride + i]
// for i < in_len (zero beyond), writes its N evaluations in blocked layout
// (the first 3N/4 when trunc).  Transforms longer than 4 times the LDS limit
// run a column pass first into tmp (count*N words of Win) with the tables ct.
template <typename Win, typename W>
hipError_t launch_ntt_fwd(const Win *in, int64_t in_stride, int64_t in_len, W *out,
                          int64_t count, int logN, int logP, const NttTune &tune,
                          const NttTab<W> &tab, W p, hipStream_t s, Win *tmp = nullptr,
                          const ColTab<W> *ct = nullptr, bool trunc = false);

// Inverse transforms (blocked input): coefficients [lo, hi) of each entry
// (clipped to N), scaled by 1/N, written to out[e*out_stride + (i - lo)] as
// Wout in [0,p).  tmp must hold count*N words of W when N exceeds the LDS
// limit; longer than 4 times that limit, the column pass uses ct.  trunc:
// the input holds the first 3N/4 evaluations of a polynomial of length
// <= 3N/4 (hi <= 3N/4).
template <typename W, typename Wout>
hipError_t launch_ntt_inv(const W *in, Wout *out, int64_t out_stride, int64_t lo, int64_t hi,
                          int64_t count, int logN, int logP, const NttTune &tune,
                          const NttTab<W> &tab, W p, W *tmp, hipStream_t s,
                          const ColTab<W> *ct = nullptr, bool trunc = false);

// ---------------------------------------------------------------------------
// Segment tables of the batched launches.  One launch serves many operand
// matrices (segments) that share a prime and a transform length; segment s
// owns the workgroups [block0, block0 + blocks) of the launch, found by a
// binary search on block0 (segments sorted by block0, no gaps).
// ---------------------------------------------------------------------------

// Forward: `count` entries, entry e at in + e*in_stride (first in_len words
// used, zero beyond), evaluations in blocked layout at out.
template <typename Win, typename W>
struct FwdSeg {
    const Win *in;
    W *out;
    int64_t in_stride, in_len, count, block0;
    int32_t trunc, pad;  // trunc: evaluations at the first 3N/4 positions only
};

// Inverse: `count` entries of blocked evaluations at in; coefficients [lo, hi)
// of entry e to out + e*out_stride.  tmp (count*N words) when N exceeds the
// LDS limit; the merge pass owns workgroups [mblock0, ...) of its own launch.
template <typename W, typename Wout>
struct InvSeg {
    const W *in;
    Wout *out;
    W *tmp;
    int64_t out_stride, lo, hi, count, block0, mblock0;
    int32_t trunc, pad;  // trunc: 3N/4 evaluations of a product of length <= 3N/4
};

// Workgroups one segment of `count` entries needs at length 2^logN (transform
// launch), and for the merge pass (0 when N fits in LDS).
template <typename W>
int64_t ntt_seg_blocks(int64_t count, int logN, bool trunc);
template <typename W>
int64_t merge_seg_blocks(int64_t count, int64_t width, int logN, bool trunc);

// segs: device array of nseg segments; blocks / mblocks: totals of the above.
// trunc: every segment of the launch is 3/4-truncated (their trunc flags set).
template <typename Win, typename W>
hipError_t launch_ntt_fwd_segs(const FwdSeg<Win, W> *segs, int nseg, int64_t blocks, int logN,
                               int logP, int log_e, const NttTab<W> &tab, W p, hipStream_t s,
                               bool trunc);
template <typename W, typename Wout>
hipError_t launch_ntt_inv_segs(const InvSeg<W, Wout> *segs, int nseg, int64_t blocks,
                               int64_t mblocks, int logN, int logP, int log_e,
                               const NttTab<W> &tab, W p, hipStream_t s, bool trunc);

enum PointwiseKind { PW_PLAIN = 0, PW_MFMA = 1 };

// C_hat[i][j][q] = sum_k A_hat[i][k][q] * B_hat[k][j][q] mod p, q < N; A_hat
// and B_hat in runs of 2^logP words, C_hat in runs of 2^logPc (logPc >= logP).
template <typename W>
hipError_t launch_pointwise(PointwiseKind kind, const W *Ah, const W *Bh, W *Ch, int m, int k,
                            int n, int64_t N, int logP, int logPc, const ModConst &mc,
                            hipStream_t s);

// Whether the int8-limb MFMA kernel supports this shape.
bool mfma_supported(int m, int k, int n, int limbs, uint64_t p);

// One product of a batched pointwise launch (evaluations in blocked layout,
// N points); tasks [block0, block0 + pw_seg_blocks(...)) are its own.
template <typename W>
struct PwSeg {
    const W *Ah, *Bh;
    W *Ch;
    int32_t m, k, n, logP;  // logP: runs of A_hat and B_hat
    int64_t N, block0;
    int32_t redc_small, logPc;  // logPc: runs of C_hat
};
int64_t pw_seg_blocks(PointwiseKind kind, int pb, int m, int n, int64_t N);
// LDS bytes of an MFMA task with pb points per workgroup
size_t pw_mfma_lds(int pb, int m, int k, int n, int limbs);
// largest points-per-workgroup (<= 2^logP) whose limb planes fit 64 KiB
int pw_pick_pb(int m, int k, int n, int limbs, int logP);
// Persistent launch over all tasks of nseg products sharing one prime:
// `grid` workgroups (0: one per task) walk the task list; MFMA tasks use pb
// points per workgroup and `lds` bytes (the maximum over the products).
template <typename W>
hipError_t launch_pointwise_segs(PointwiseKind kind, int pb, const PwSeg<W> *segs, int nseg,
                                 int64_t blocks, size_t lds, int grid, const ModConst &mc,
                                 hipStream_t s);

}  // namespace pmlg
