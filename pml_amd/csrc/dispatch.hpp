// Per-product configurations of the batched path and the dispatch between
// them (threshold rule, autotuner cache, learned cost model).
#pragma once

#include "ctx.hpp"

namespace pmlg {

struct ProdConfig {
    int school;          // 1: direct convolution, 0: transforms + pointwise products
    PointwiseKind kind;  // pointwise kernel
    int fwd_e, inv_e;    // log2 elements per thread of the register transforms
    int pb_shift;        // MFMA points per workgroup: the largest that fits, >> pb_shift
};

constexpr int kNumConfigs = 8;
extern const ProdConfig kConfigs[kNumConfigs];

// direct convolution is offered when the shorter operand has at most this
// many coefficients, chosen by the threshold rule up to kSchoolThreshold, and
// timed by the autotuner while m n k lenC min(lenA, lenB) <= kSchoolTimeCap
constexpr int64_t kSchoolMaxLen = 256;
constexpr int64_t kSchoolThreshold = 8;
// most identical products the autotuner times together per configuration
constexpr int kTuneCopies = 16;
constexpr double kSchoolTimeCap = 8.6e9;

bool config_valid(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB, int c);
int threshold_config(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB);
int choose_config(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                  int method);
// configuration picked by the learned cost model, or -1 when none is loaded;
// pred (if set) receives the predicted log(ms) of every configuration
int learned_config(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                   float *pred = nullptr);
// The model's top candidates whose predicted times lie within kRefineMargin
// of its best are timed once per shape class (measured refinement).  The
// network's error is about 24% RMS in log time and not uniform across
// configurations (smooth features cannot show the MFMA tiles' padding steps
// at multiples of 32), so the margin is wide: on 2,000 held-out shapes, with
// the recorded times standing in for the live ones, 15% brings the total
// within 0.05% of the per-shape best with 4 candidates a shape (5%: 0.3%
// above it with 2.6).
constexpr double kRefineMargin = 0.15;
bool refine_enabled();

pmlg_status build_plan(pmlg_batch_plan *out, const pmlg_mul_desc *descs, int64_t count, int flags,
                       void *ws, size_t ws_bytes, size_t *need_bytes);
pmlg_status run_plan(pmlg_batch_plan plan, hipStream_t s);
// Times the configurations in `mask` (bit c: configuration c) for one shape;
// with the full mask the winner also goes to the autotuner cache.
constexpr unsigned kAllConfigs = (1u << kNumConfigs) - 1;
pmlg_status time_configs(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                         int reps, double *ms, int *best, unsigned mask = kAllConfigs);

}  // namespace pmlg
