// Batched products: many independent C_i = A_i * B_i of uneven shapes,
// lengths and moduli, planned once and launched as one HIP graph.
//
// A plan groups the products by what a launch can share:
//   * forward transforms by (modulus, transform length, radix): one
//     segment-table launch per group covers every entry of every operand;
//   * pointwise products by (modulus, kernel, points per workgroup): one
//     persistent launch per group walks the task list of all its products;
//   * inverse transforms by (modulus, length, radix, output scaling);
//   * short products by direct (schoolbook) convolution, one persistent
//     launch per modulus.
// The launch sequence is captured into a hipGraph on the first run and
// replayed afterwards.  Each product's configuration (algorithm, pointwise
// kernel, radix of the register transforms, points per MFMA workgroup) comes
// from the caller, the learned cost model, the autotuner's cache or the
// threshold rule, in that order; the result is the same exact product
// whichever runs.
#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <vector>

#include "ctx.hpp"
#include "dispatch.hpp"
#include "host_arith.hpp"

namespace pmlg {

pmlg_status get_tables(pmlg_ctx ctx, int logN, const Tables **out);
template <typename W>
const NttTab<W> &tab_of(const Tables &T);

namespace {

// ---------------------------------------------------------------------------
// direct convolution for short operands
// ---------------------------------------------------------------------------
struct SchoolSeg {
    const uint64_t *A, *B;
    uint64_t *C;
    int64_t m, k, n, lenA, lenB, block0;
};

constexpr int kSchoolThreads = 256;

__device__ __forceinline__ const SchoolSeg &school_seg_of(const SchoolSeg *segs, int nseg, int64_t b)
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

// task = (output entry, 256 consecutive coefficients); each lane accumulates
// C[i][j][d] = sum_kk sum_t A[i][kk][t] B[kk][j][d-t] exactly in 128 bits,
// folding every 15 products (each < 2^124) so the sum stays below 2^128.
__global__ void __launch_bounds__(kSchoolThreads)
school_segs(const SchoolSeg *__restrict__ segs, int nseg, int64_t blocks, ModConst mc)
{
    for (int64_t b = blockIdx.x; b < blocks; b += gridDim.x) {
        const SchoolSeg &sg = school_seg_of(segs, nseg, b);
        const int64_t lenC = sg.lenA + sg.lenB - 1;
        const int64_t chunks = (lenC + kSchoolThreads - 1) / kSchoolThreads;
        const int64_t lb = b - sg.block0;
        const int64_t e = lb / chunks;
        const int64_t d = (lb - e * chunks) * kSchoolThreads + threadIdx.x;
        if (d >= lenC)
            continue;
        const int64_t i = e / sg.n, j = e - i * sg.n;
        const int64_t t0 = d - sg.lenB + 1 > 0 ? d - sg.lenB + 1 : 0;
        const int64_t t1 = d < sg.lenA - 1 ? d : sg.lenA - 1;
        uint64_t hi = 0, lo = 0;
        int cnt = 0;
        for (int64_t kk = 0; kk < sg.k; ++kk) {
            const uint64_t *a = sg.A + (i * sg.k + kk) * sg.lenA;
            const uint64_t *bb = sg.B + (kk * sg.n + j) * sg.lenB + d;
            for (int64_t t = t0; t <= t1; ++t) {
                mac128(hi, lo, a[t], bb[-t]);
                if (++cnt == 15) {
                    lo = reduce128(hi, lo, mc);
                    hi = 0;
                    cnt = 0;
                }
            }
        }
        sg.C[e * lenC + d] = reduce128(hi, lo, mc);
    }
}

struct FillSeg {
    uint64_t *ptr;
    int64_t len;
};

// zero outputs of products with an empty inner dimension
__global__ void fill_zero_segs(const FillSeg *__restrict__ segs, int nseg)
{
    for (int s = 0; s < nseg; ++s)
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < segs[s].len;
             i += (int64_t)gridDim.x * blockDim.x)
            segs[s].ptr[i] = 0;
}

// splitmix64 stream reduced to [0, p) (inputs of the autotuner's timing runs)
__global__ void fill_random_kernel(uint64_t *__restrict__ out, int64_t n, uint64_t p, uint64_t seed)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + (uint64_t)(i + 1) * 0x9e3779b97f4a7c15ull;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        z ^= z >> 31;
        out[i] = (uint64_t)(((unsigned __int128)z * p) >> 64);
    }
}

unsigned grid_for(int64_t items, int64_t cap = 4096)
{
    int64_t g = (items + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(g, cap));
}

}  // namespace

// ---------------------------------------------------------------------------
// configurations
// ---------------------------------------------------------------------------
const ProdConfig kConfigs[kNumConfigs] = {
    // school, kind,     fwd_e, inv_e, pb_shift
    {1, PW_PLAIN, 5, 4, 0},  // 0: direct convolution
    {0, PW_PLAIN, 5, 4, 0},  // 1: transforms + plain pointwise
    {0, PW_MFMA, 5, 4, 0},   // 2: transforms + int8-limb MFMA pointwise
    {0, PW_PLAIN, 4, 4, 0},  // 3: plain, radix-16 forward passes
    {0, PW_MFMA, 4, 4, 0},   // 4: MFMA, radix-16 forward passes
    {0, PW_MFMA, 5, 5, 0},   // 5: MFMA, radix-32 inverse passes
    {0, PW_MFMA, 5, 4, 1},   // 6: MFMA, half the points per workgroup
    {0, PW_PLAIN, 5, 5, 0},  // 7: plain, radix-32 inverse passes
};

bool config_valid(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                  int c)
{
    if (c < 0 || c >= kNumConfigs)
        return false;
    const ProdConfig &cf = kConfigs[c];
    if (cf.school)
        return std::min(lenA, lenB) <= kSchoolMaxLen;
    // the segment-table transforms go up to 4 times the LDS limit
    const int seg_max = ctx->small ? NttLimits<uint32_t>::max_log_m + 2
                                   : NttLimits<uint64_t>::max_log_m + 2;
    if (lenA > 0 && lenB > 0 &&
        ceil_log2(lenA + lenB - 1) > std::min(ctx->max_log, seg_max))
        return false;
    if (cf.kind == PW_MFMA)
        return ctx->mc.limbs > 0 && mfma_supported((int)m, (int)k, (int)n, ctx->mc.limbs, ctx->p);
    return true;
}

// Threshold rule (the fallback, and the baseline the learned choice is
// measured against), banded in the style of the reference's dispatch
// (mul.c:57-96: bands over the dimension and the length): the word size, the
// shorter operand's length (<= 8, 32, 128, 512, 2048, longer), the largest
// dimension (<= 8, 16, 32, 64) and the product length (< 2^11, 2^13, 2^15,
// longer) select a cell, whose configuration was fitted on the recorded
// timings (bands.inc, generated by `python -m pml_amd.learn bands`; the same
// cell function is pml_amd.learn.band_index).  A cell without a fit, or whose
// configuration cannot run the shape, falls back to the fixed rule: direct
// convolution when both operands are short, else transforms with the MFMA
// pointwise kernel whenever the shape and modulus allow it, else plain.
#include "bands.inc"

int band_index(bool small, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB)
{
    static const int64_t kShort[5] = {8, 32, 128, 512, 2048};
    static const int64_t kDim[3] = {8, 16, 32};
    const int w = small ? 0 : 1;
    const int64_t sh = std::min(lenA, lenB);
    int sb = 5;
    for (int i = 0; i < 5; ++i)
        if (sh <= kShort[i]) {
            sb = i;
            break;
        }
    const int64_t dim = std::max(m, std::max(k, n));
    int db = 3;
    for (int i = 0; i < 3; ++i)
        if (dim <= kDim[i]) {
            db = i;
            break;
        }
    const int lg = ceil_log2(lenA + lenB - 1);
    const int lb = std::min(3, std::max(0, (lg - 9) / 2));
    return ((w * 6 + sb) * 4 + db) * 4 + lb;
}

int threshold_config(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB)
{
    if (lenA > 0 && lenB > 0) {
        const int c = kBandConfig[band_index(ctx->small, m, k, n, lenA, lenB)];
        if (c >= 0 && config_valid(ctx, m, k, n, lenA, lenB, c))
            return c;
    }
    if (std::min(lenA, lenB) <= kSchoolThreshold && config_valid(ctx, m, k, n, lenA, lenB, 0))
        return 0;
    if (config_valid(ctx, m, k, n, lenA, lenB, 2))
        return 2;
    return 1;
}

static uint64_t tune_key(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB)
{
    // shape class: log2 of each dimension and of both lengths
    auto lg = [](int64_t x) { return (uint64_t)ceil_log2(std::max<int64_t>(x, 1)); };
    return (uint64_t)(ctx->small ? 1 : 0) | lg(m) << 1 | lg(k) << 5 | lg(n) << 9 | lg(lenA) << 13 |
           lg(lenB) << 18;
}

int choose_config(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                  int method)
{
    if (method == PMLG_METHOD_PLAIN)
        return 1;
    if (method == PMLG_METHOD_MFMA)
        return 2;
    if (method == PMLG_METHOD_THRESHOLD)
        return threshold_config(ctx, m, k, n, lenA, lenB);
    float pred[kNumConfigs];
    int c = learned_config(ctx, m, k, n, lenA, lenB, pred);
    if (c >= 0 && config_valid(ctx, m, k, n, lenA, lenB, c)) {
        // candidates the model cannot separate: predicted within kRefineMargin of its best
        unsigned mask = 0;
        int ncand = 0;
        const float bound = pred[c] + (float)std::log1p(kRefineMargin);
        for (int cc = 0; cc < kNumConfigs; ++cc)
            if (pred[cc] <= bound) {
                mask |= 1u << cc;
                ++ncand;
            }
        if (ncand <= 1 || !refine_enabled())
            return c;
        const uint64_t key = tune_key(ctx, m, k, n, lenA, lenB);
        {
            std::lock_guard<std::mutex> lk(ctx->tune_mu);
            auto it = ctx->refined.find(key);
            if (it != ctx->refined.end() && config_valid(ctx, m, k, n, lenA, lenB, it->second))
                return it->second;
        }
        double ms[kNumConfigs];
        int best = -1;
        if (time_configs(ctx, m, k, n, lenA, lenB, 1, ms, &best, mask) == PMLG_OK && best >= 0) {
            std::lock_guard<std::mutex> lk(ctx->tune_mu);
            ctx->refined[key] = best;
            return best;
        }
        return c;
    }
    {
        std::lock_guard<std::mutex> lk(ctx->tune_mu);
        auto it = ctx->tuned.find(tune_key(ctx, m, k, n, lenA, lenB));
        if (it != ctx->tuned.end() && config_valid(ctx, m, k, n, lenA, lenB, it->second))
            return it->second;
    }
    return threshold_config(ctx, m, k, n, lenA, lenB);
}

}  // namespace pmlg

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
using namespace pmlg;

namespace {

struct FwdGroup {
    pmlg_ctx ctx;
    int logN, logP, fwd_e;
    size_t seg_off;  // byte offset of the segment array in the plan's table block
    int nseg;
    int64_t blocks;
    bool trunc;  // 3/4-truncated products (their own launches: the split may differ)
};
struct InvGroup {
    pmlg_ctx ctx;
    int logN, logP, inv_e;
    bool redc;
    size_t seg_off;
    int nseg;
    int64_t blocks, mblocks;
    bool trunc;
};
struct PwGroup {
    pmlg_ctx ctx;
    PointwiseKind kind;
    int pb;
    size_t seg_off;
    int nseg;
    int64_t blocks;
    size_t lds;
};
struct SchoolGroup {
    pmlg_ctx ctx;
    size_t seg_off;
    int nseg;
    int64_t blocks;
};

struct ProdInfo {
    int config;
    int logN;
    size_t ah, bh, ch, tmp;  // word offsets in the arena (bytes / word size)
};

}  // namespace

struct pmlg_batch_plan_s {
    int device = 0;
    int flags = 0;
    hipStream_t stream = nullptr;
    void *arena = nullptr;  // evaluations of every product
    size_t arena_bytes = 0;
    void *tables = nullptr;  // segment arrays of every launch
    size_t table_bytes = 0;
    bool own_mem = true;
    std::vector<FwdGroup> fwd;
    std::vector<InvGroup> inv;
    std::vector<PwGroup> pw;
    std::vector<SchoolGroup> school;
    int nfill = 0;
    size_t fill_off = 0;
    std::vector<ProdInfo> prods;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // stage timing of graph replays (pmlg_batch_plan_time_stages): one graph
    // per stage (rgraph, rexec), launched in turn with events recorded
    // between them; `ring` sets of events used in turn, replays[r] counting
    // the replays of set r since the last read
    int ring = 0;
    std::vector<hipGraph_t> rgraph;
    std::vector<hipGraphExec_t> rexec;
    std::vector<std::array<hipEvent_t, 4>> rev;
    std::vector<int64_t> replays;
    int64_t rnext = 0;
};

namespace {

struct Blob {
    std::vector<unsigned char> bytes;
    template <typename T>
    size_t add(const std::vector<T> &v)
    {
        size_t off = (bytes.size() + 255) & ~(size_t)255;
        bytes.resize(off + v.size() * sizeof(T));
        if (!v.empty())
            std::memcpy(bytes.data() + off, v.data(), v.size() * sizeof(T));
        return off;
    }
};

template <typename W>
const NttTab<W> *tables_for(pmlg_ctx ctx, int logN, pmlg_status *st)
{
    const Tables *T = nullptr;
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    *st = get_tables(ctx, logN, &T);
    return *st == PMLG_OK ? &tab_of<W>(*T) : nullptr;
}

template <typename W>
pmlg_status launch_groups(pmlg_batch_plan plan, hipStream_t s)
{
    const unsigned char *tb = static_cast<const unsigned char *>(plan->tables);
    for (const FwdGroup &g : plan->fwd) {
        if ((g.ctx->small ? 4 : 8) != sizeof(W))
            continue;
        pmlg_status st;
        const NttTab<W> *tab = tables_for<W>(g.ctx, g.logN, &st);
        if (!tab)
            return st;
        HIP_TRY((launch_ntt_fwd_segs<uint64_t, W>(
                    reinterpret_cast<const FwdSeg<uint64_t, W> *>(tb + g.seg_off), g.nseg,
                    g.blocks, g.logN, g.logP, g.fwd_e, *tab, (W)g.ctx->p, s, g.trunc)),
                "batched forward transforms");
    }
    return PMLG_OK;
}

template <typename W>
pmlg_status launch_pw(pmlg_batch_plan plan, hipStream_t s)
{
    const unsigned char *tb = static_cast<const unsigned char *>(plan->tables);
    for (const PwGroup &g : plan->pw) {
        if ((g.ctx->small ? 4 : 8) != sizeof(W))
            continue;
        // persistent grid: a few workgroups per CU, never more than the tasks
        const int grid = (int)std::min<int64_t>(g.blocks, g.kind == PW_MFMA ? 512 : 2048);
        HIP_TRY(launch_pointwise_segs<W>(g.kind, g.pb,
                                         reinterpret_cast<const PwSeg<W> *>(tb + g.seg_off),
                                         g.nseg, g.blocks, g.lds, grid, g.ctx->mc, s),
                "batched pointwise products");
    }
    return PMLG_OK;
}

template <typename W>
pmlg_status launch_inv(pmlg_batch_plan plan, hipStream_t s)
{
    const unsigned char *tb = static_cast<const unsigned char *>(plan->tables);
    for (const InvGroup &g : plan->inv) {
        if ((g.ctx->small ? 4 : 8) != sizeof(W))
            continue;
        pmlg_status st;
        const NttTab<W> *tab = tables_for<W>(g.ctx, g.logN, &st);
        if (!tab)
            return st;
        NttTab<W> itab = *tab;
        if (g.redc) {
            itab.ninv = tab->ninvr;
            itab.ninvq = tab->ninvrq;
            itab.ninv2 = tab->ninvr2;
            itab.ninv2q = tab->ninvr2q;
        }
        HIP_TRY((launch_ntt_inv_segs<W, uint64_t>(
                    reinterpret_cast<const InvSeg<W, uint64_t> *>(tb + g.seg_off), g.nseg,
                    g.blocks, g.mblocks, g.logN, g.logP, g.inv_e, itab, (W)g.ctx->p, s, g.trunc)),
                "batched inverse transforms");
    }
    return PMLG_OK;
}

// stage timing of an eager run: events between the stages (forward
// transforms, direct convolutions + fills + pointwise, inverse transforms)
struct StageEvents {
    hipEvent_t ev[4];
    bool on;
};

// `only` in 0..2 launches that stage alone (the per-stage graphs of
// pmlg_batch_plan_time_stages), -1 all three
pmlg_status launch_all(pmlg_batch_plan plan, hipStream_t s, StageEvents *se = nullptr,
                       int only = -1)
{
    const unsigned char *tb = static_cast<const unsigned char *>(plan->tables);
    pmlg_status st;
    if (se)
        HIP_TRY(hipEventRecord(se->ev[0], s), "hipEventRecord");
    if (only < 0 || only == 0)
        if ((st = launch_groups<uint64_t>(plan, s)) != PMLG_OK ||
            (st = launch_groups<uint32_t>(plan, s)) != PMLG_OK)
            return st;
    if (se)
        HIP_TRY(hipEventRecord(se->ev[1], s), "hipEventRecord");
    if (only >= 0 && only != 1)
        goto inverse;
    for (const SchoolGroup &g : plan->school) {
        const int grid = (int)std::min<int64_t>(g.blocks, 2048);
        hipLaunchKernelGGL(school_segs, dim3(grid), dim3(kSchoolThreads), 0, s,
                           reinterpret_cast<const SchoolSeg *>(tb + g.seg_off), g.nseg, g.blocks,
                           g.ctx->mc);
        HIP_TRY(hipGetLastError(), "batched direct convolutions");
    }
    if (plan->nfill > 0) {
        hipLaunchKernelGGL(fill_zero_segs, dim3(256), dim3(256), 0, s,
                           reinterpret_cast<const FillSeg *>(tb + plan->fill_off), plan->nfill);
        HIP_TRY(hipGetLastError(), "zero fill");
    }
    if ((st = launch_pw<uint64_t>(plan, s)) != PMLG_OK || (st = launch_pw<uint32_t>(plan, s)) != PMLG_OK)
        return st;
inverse:
    if (se)
        HIP_TRY(hipEventRecord(se->ev[2], s), "hipEventRecord");
    if (only < 0 || only == 2)
        if ((st = launch_inv<uint64_t>(plan, s)) != PMLG_OK ||
            (st = launch_inv<uint32_t>(plan, s)) != PMLG_OK)
            return st;
    if (se)
        HIP_TRY(hipEventRecord(se->ev[3], s), "hipEventRecord");
    return PMLG_OK;
}

template <typename W>
void add_segments(const pmlg_mul_desc &d, const ProdInfo &pi, unsigned char *arena,
                  std::map<std::tuple<pmlg_ctx, int, int, int>, std::vector<FwdSeg<uint64_t, W>>> &fwd,
                  std::map<std::tuple<pmlg_ctx, int, int, int>, std::vector<InvSeg<W, uint64_t>>> &inv,
                  std::map<std::tuple<pmlg_ctx, int, int>, std::vector<PwSeg<W>>> &pw,
                  std::map<std::tuple<pmlg_ctx, int, int>, size_t> &pw_lds)
{
    pmlg_ctx ctx = d.ctx;
    const ProdConfig &cf = kConfigs[pi.config];
    const int logN = pi.logN;
    const int64_t N = (int64_t)1 << logN;
    const int logP = eval_log_run_in<W>(logN);
    W *Ah = reinterpret_cast<W *>(arena) + pi.ah;
    W *Bh = reinterpret_cast<W *>(arena) + pi.bh;
    W *Ch = reinterpret_cast<W *>(arena) + pi.ch;
    W *tmp = pi.tmp != (size_t)-1 ? reinterpret_cast<W *>(arena) + pi.tmp : nullptr;
    const int64_t lenC = d.lenA + d.lenB - 1;
    // a product of length <= 3N/4 needs 3N/4 evaluations only
    const int trunc = trunc_ok<W>(logN, lenC) ? 1 : 0;
    auto &fv = fwd[{ctx, logN, cf.fwd_e, trunc}];
    fv.push_back(FwdSeg<uint64_t, W>{d.A, Ah, d.lenA, d.lenA, d.m * d.k, 0, trunc, 0});
    fv.push_back(FwdSeg<uint64_t, W>{d.B, Bh, d.lenB, d.lenB, d.k * d.n, 0, trunc, 0});
    const bool redc = cf.kind == PW_MFMA && sizeof(W) == 8;
    inv[{ctx, logN, cf.inv_e, (redc ? 1 : 0) | (trunc << 1)}].push_back(
        InvSeg<W, uint64_t>{Ch, d.C, tmp, lenC, 0, lenC, d.m * d.n, 0, 0, trunc, 0});
    int pb = 0;
    if (cf.kind == PW_MFMA) {
        pb = pw_pick_pb((int)d.m, (int)d.k, (int)d.n, ctx->mc.limbs, logP) >> cf.pb_shift;
        pb = std::max(pb, 1);
    }
    PwSeg<W> ps{};
    ps.Ah = Ah;
    ps.Bh = Bh;
    ps.Ch = Ch;
    ps.m = (int32_t)d.m;
    ps.k = (int32_t)d.k;
    ps.n = (int32_t)d.n;
    ps.logP = logP;
    ps.logPc = eval_log_run_out<W>(logN);
    ps.N = trunc ? 3 * (N >> 2) : N;
    ps.redc_small = 0;
    if (redc) {
        const unsigned __int128 vmax = (unsigned __int128)d.k * (ctx->p - 1) * (ctx->p - 1);
        ps.redc_small = vmax < (((unsigned __int128)(3 * ctx->p - 1)) << 64) ? 1 : 0;
    }
    const auto key = std::make_tuple(ctx, (int)cf.kind, pb);
    pw[key].push_back(ps);
    if (cf.kind == PW_MFMA) {
        size_t &l = pw_lds[key];
        l = std::max(l, pw_mfma_lds(pb, (int)d.m, (int)d.k, (int)d.n, ctx->mc.limbs));
    }
}

template <typename W>
void finish_groups(pmlg_batch_plan plan, Blob &blob,
                   std::map<std::tuple<pmlg_ctx, int, int, int>, std::vector<FwdSeg<uint64_t, W>>> &fwd,
                   std::map<std::tuple<pmlg_ctx, int, int, int>, std::vector<InvSeg<W, uint64_t>>> &inv,
                   std::map<std::tuple<pmlg_ctx, int, int>, std::vector<PwSeg<W>>> &pw,
                   std::map<std::tuple<pmlg_ctx, int, int>, size_t> &pw_lds)
{
    for (auto &kv : fwd) {
        const int logN = std::get<1>(kv.first);
        const bool trunc = std::get<3>(kv.first) != 0;
        int64_t b = 0;
        for (auto &sg : kv.second) {
            sg.block0 = b;
            b += ntt_seg_blocks<W>(sg.count, logN, trunc);
        }
        FwdGroup g{std::get<0>(kv.first), logN, eval_log_run_in<W>(logN),
                   std::get<2>(kv.first), blob.add(kv.second), (int)kv.second.size(), b, trunc};
        plan->fwd.push_back(g);
    }
    for (auto &kv : inv) {
        const int logN = std::get<1>(kv.first);
        const bool trunc = (std::get<3>(kv.first) & 2) != 0;
        int64_t b = 0, mb = 0;
        for (auto &sg : kv.second) {
            sg.block0 = b;
            sg.mblock0 = mb;
            b += ntt_seg_blocks<W>(sg.count, logN, trunc);
            mb += merge_seg_blocks<W>(sg.count, sg.hi - sg.lo, logN, trunc);
        }
        InvGroup g{std::get<0>(kv.first), logN, eval_log_run_out<W>(logN),
                   std::get<2>(kv.first), (std::get<3>(kv.first) & 1) != 0, blob.add(kv.second),
                   (int)kv.second.size(), b, mb, trunc};
        plan->inv.push_back(g);
    }
    for (auto &kv : pw) {
        const PointwiseKind kind = (PointwiseKind)std::get<1>(kv.first);
        const int pb = std::get<2>(kv.first);
        int64_t b = 0;
        for (auto &sg : kv.second) {
            sg.block0 = b;
            b += pw_seg_blocks(kind, pb, sg.m, sg.n, sg.N);
        }
        PwGroup g{std::get<0>(kv.first), kind, pb, blob.add(kv.second), (int)kv.second.size(), b,
                  kind == PW_MFMA ? pw_lds[kv.first] : 0};
        plan->pw.push_back(g);
    }
}

void free_ring(pmlg_batch_plan plan)
{
    for (hipGraphExec_t e : plan->rexec)
        if (e)
            (void)hipGraphExecDestroy(e);
    for (hipGraph_t g : plan->rgraph)
        if (g)
            (void)hipGraphDestroy(g);
    for (auto &ev : plan->rev)
        for (hipEvent_t e : ev)
            if (e)
                (void)hipEventDestroy(e);
    plan->rexec.clear();
    plan->rgraph.clear();
    plan->rev.clear();
    plan->replays.clear();
    plan->rnext = 0;
}

void free_plan_memory(pmlg_batch_plan plan)
{
    DeviceGuard g(plan->device);
    if (plan->exec)
        (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph)
        (void)hipGraphDestroy(plan->graph);
    plan->exec = nullptr;
    plan->graph = nullptr;
    if (plan->stream)
        (void)hipStreamSynchronize(plan->stream);
    free_ring(plan);
    if (plan->own_mem) {
        if (plan->arena)
            (void)hipFree(plan->arena);
        if (plan->tables)
            (void)hipFree(plan->tables);
    }
    plan->arena = plan->tables = nullptr;
    if (plan->stream)
        (void)hipStreamDestroy(plan->stream);
    plan->stream = nullptr;
}

}  // namespace

namespace pmlg {

// Builds a plan; with `ws` non-null the arena and tables are carved from it
// (ws_bytes long) instead of being allocated.
pmlg_status build_plan(pmlg_batch_plan *out, const pmlg_mul_desc *descs, int64_t count, int flags,
                       void *ws, size_t ws_bytes, size_t *need_bytes)
{
    if (!out || (count > 0 && !descs) || count < 0)
        return fail(PMLG_ERR_ARG, "batch: bad arguments");
    *out = nullptr;
    struct PlanDel {
        void operator()(pmlg_batch_plan_s *p) const
        {
            free_plan_memory(p);
            delete p;
        }
    };
    std::unique_ptr<pmlg_batch_plan_s, PlanDel> plan(new (std::nothrow) pmlg_batch_plan_s());
    if (!plan)
        return fail(PMLG_ERR_ALLOC, "out of host memory");
    plan->flags = flags;
    plan->device = count > 0 && descs[0].ctx ? descs[0].ctx->device : 0;
    plan->prods.resize(count);

    // configurations and arena layout
    size_t words_bytes = 0;  // running byte offset, aligned per buffer
    auto take = [&](size_t bytes) {
        size_t off = (words_bytes + 255) & ~(size_t)255;
        words_bytes = off + bytes;
        return off;
    };
    std::map<pmlg_ctx, std::vector<SchoolSeg>> school;
    std::vector<FillSeg> fills;
    for (int64_t i = 0; i < count; ++i) {
        const pmlg_mul_desc &d = descs[i];
        pmlg_status st = check_ctx(d.ctx);
        if (st != PMLG_OK)
            return st;
        if (d.ctx->device != plan->device)
            return fail(PMLG_ERR_ARG, "batch: all contexts must use one device");
        if (d.m < 0 || d.k < 0 || d.n < 0 || d.lenA < 0 || d.lenB < 0 || d.m > 4096 ||
            d.k > 4096 || d.n > 4096)
            return fail(PMLG_ERR_ARG, "batch: bad dimension or length in product " + std::to_string(i));
        ProdInfo &pi = plan->prods[i];
        pi.config = -1;
        pi.logN = 0;
        pi.tmp = (size_t)-1;
        if (d.lenA == 0 || d.lenB == 0 || d.m == 0 || d.n == 0)
            continue;  // nothing to write
        if (!d.C || (d.k > 0 && (!d.A || !d.B)))
            return fail(PMLG_ERR_ARG, "batch: null buffer in product " + std::to_string(i));
        const int64_t lenC = d.lenA + d.lenB - 1;
        if (d.k == 0) {
            fills.push_back(FillSeg{d.C, d.m * d.n * lenC});
            continue;
        }
        int c = d.config >= 0 ? d.config : choose_config(d.ctx, d.m, d.k, d.n, d.lenA, d.lenB, d.method);
        if (!config_valid(d.ctx, d.m, d.k, d.n, d.lenA, d.lenB, c))
            return fail(PMLG_ERR_UNSUPPORTED,
                        "batch: configuration " + std::to_string(c) + " cannot run product " +
                            std::to_string(i));
        pi.config = c;
        if (kConfigs[c].school) {
            school[d.ctx].push_back(SchoolSeg{d.A, d.B, d.C, d.m, d.k, d.n, d.lenA, d.lenB, 0});
            continue;
        }
        pi.logN = ceil_log2(lenC);
        const size_t w = d.ctx->small ? 4 : 8;
        const int64_t N = (int64_t)1 << pi.logN;
        pi.ah = take((size_t)(d.m * d.k * N) * w) / w;
        pi.bh = take((size_t)(d.k * d.n * N) * w) / w;
        pi.ch = take((size_t)(d.m * d.n * N) * w) / w;
        const int lim = d.ctx->small ? NttLimits<uint32_t>::max_log_m : NttLimits<uint64_t>::max_log_m;
        if (pi.logN > lim)
            pi.tmp = take((size_t)(d.m * d.n * N) * w) / w;
    }
    const size_t arena_bytes = std::max<size_t>(words_bytes, 256);

    // segment tables (device addresses depend on the arena, so place it first)
    unsigned char *arena = nullptr;
    size_t table_cap = 0;
    if (ws) {
        if (arena_bytes > ws_bytes)
            return fail(PMLG_ERR_ARG, "batch: workspace too small");
        arena = static_cast<unsigned char *>(ws);
    } else {
        DeviceGuard g(plan->device);
        HIP_TRY(hipMalloc(&plan->arena, arena_bytes), "hipMalloc(batch arena)");
        arena = static_cast<unsigned char *>(plan->arena);
    }
    plan->arena_bytes = arena_bytes;

    Blob blob;
    std::map<std::tuple<pmlg_ctx, int, int, int>, std::vector<FwdSeg<uint64_t, uint64_t>>> f64;
    std::map<std::tuple<pmlg_ctx, int, int, int>, std::vector<FwdSeg<uint64_t, uint32_t>>> f32;
    std::map<std::tuple<pmlg_ctx, int, int, int>, std::vector<InvSeg<uint64_t, uint64_t>>> i64;
    std::map<std::tuple<pmlg_ctx, int, int, int>, std::vector<InvSeg<uint32_t, uint64_t>>> i32;
    std::map<std::tuple<pmlg_ctx, int, int>, std::vector<PwSeg<uint64_t>>> p64;
    std::map<std::tuple<pmlg_ctx, int, int>, std::vector<PwSeg<uint32_t>>> p32;
    std::map<std::tuple<pmlg_ctx, int, int>, size_t> l64, l32;
    for (int64_t i = 0; i < count; ++i) {
        const ProdInfo &pi = plan->prods[i];
        if (pi.config < 0 || kConfigs[pi.config].school)
            continue;
        if (descs[i].ctx->small)
            add_segments<uint32_t>(descs[i], pi, arena, f32, i32, p32, l32);
        else
            add_segments<uint64_t>(descs[i], pi, arena, f64, i64, p64, l64);
    }
    finish_groups<uint64_t>(plan.get(), blob, f64, i64, p64, l64);
    finish_groups<uint32_t>(plan.get(), blob, f32, i32, p32, l32);
    for (auto &kv : school) {
        int64_t b = 0;
        for (auto &sg : kv.second) {
            sg.block0 = b;
            const int64_t lenC = sg.lenA + sg.lenB - 1;
            b += sg.m * sg.n * ((lenC + kSchoolThreads - 1) / kSchoolThreads);
        }
        plan->school.push_back(SchoolGroup{kv.first, blob.add(kv.second), (int)kv.second.size(), b});
    }
    plan->nfill = (int)fills.size();
    if (plan->nfill)
        plan->fill_off = blob.add(fills);
    table_cap = std::max<size_t>(blob.bytes.size(), 256);
    if (need_bytes)
        *need_bytes = ((arena_bytes + 255) & ~(size_t)255) + table_cap;
    DeviceGuard g(plan->device);
    if (ws) {
        const size_t toff = (arena_bytes + 255) & ~(size_t)255;
        plan->own_mem = false;
        if (toff + table_cap > ws_bytes)
            return fail(PMLG_ERR_ARG, "batch: workspace too small");
        plan->tables = static_cast<unsigned char *>(ws) + toff;
    } else {
        HIP_TRY(hipMalloc(&plan->tables, table_cap), "hipMalloc(batch tables)");
    }
    plan->table_bytes = table_cap;
    if (!blob.bytes.empty()) {
        HIP_TRY(hipMemcpy(plan->tables, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice),
                "hipMemcpy(batch tables)");
    }
    HIP_TRY(hipStreamCreateWithFlags(&plan->stream, hipStreamNonBlocking), "hipStreamCreate");
    *out = plan.release();
    return PMLG_OK;
}

pmlg_status run_plan(pmlg_batch_plan plan, hipStream_t s)
{
    DeviceGuard g(plan->device);
    if (!(plan->flags & PMLG_BATCH_GRAPH))
        return launch_all(plan, s);
    if (!plan->exec) {
        // One eager pass first: it builds the twiddle tables and sets the
        // kernels' LDS attributes (host work that must not happen while the
        // stream is being captured); its results are those of the graph.
        pmlg_status st0 = launch_all(plan, s);
        if (st0 != PMLG_OK)
            return st0;
        HIP_TRY(hipStreamSynchronize(s), "hipStreamSynchronize");
        HIP_TRY(hipStreamBeginCapture(plan->stream, hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture");
        pmlg_status st = launch_all(plan, plan->stream);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(plan->stream, &graph);
        if (st != PMLG_OK) {
            if (graph)
                (void)hipGraphDestroy(graph);
            return st;
        }
        HIP_TRY(e, "hipStreamEndCapture");
        plan->graph = graph;
        HIP_TRY(hipGraphInstantiate(&plan->exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        // one graph per stage, for timing the replays with events between them
        if (plan->ring > 0) {
            for (int stg = 0; stg < 3; ++stg) {
                HIP_TRY(hipStreamBeginCapture(plan->stream, hipStreamCaptureModeThreadLocal),
                        "hipStreamBeginCapture");
                pmlg_status st2 = launch_all(plan, plan->stream, nullptr, stg);
                hipGraph_t g2 = nullptr;
                hipError_t e2 = hipStreamEndCapture(plan->stream, &g2);
                if (st2 != PMLG_OK) {
                    if (g2)
                        (void)hipGraphDestroy(g2);
                    return st2;
                }
                HIP_TRY(e2, "hipStreamEndCapture");
                plan->rgraph.push_back(g2);
                hipGraphExec_t x2 = nullptr;
                HIP_TRY(hipGraphInstantiate(&x2, g2, nullptr, nullptr, 0), "hipGraphInstantiate");
                plan->rexec.push_back(x2);
            }
            for (int r = 0; r < plan->ring; ++r) {
                std::array<hipEvent_t, 4> ev{};
                for (hipEvent_t &x : ev)
                    HIP_TRY(hipEventCreate(&x), "hipEventCreate");
                plan->rev.push_back(ev);
                plan->replays.push_back(0);
            }
        }
        return PMLG_OK;  // this run's results come from the eager pass
    }
    if (!plan->rexec.empty()) {
        // the three stage graphs with events recorded between them
        const size_t r = (size_t)(plan->rnext++ % (int64_t)plan->rev.size());
        const std::array<hipEvent_t, 4> &ev = plan->rev[r];
        for (int stg = 0; stg < 3; ++stg) {
            HIP_TRY(hipEventRecord(ev[stg], s), "hipEventRecord");
            HIP_TRY(hipGraphLaunch(plan->rexec[stg], s), "hipGraphLaunch");
        }
        HIP_TRY(hipEventRecord(ev[3], s), "hipEventRecord");
        plan->replays[r]++;
        return PMLG_OK;
    }
    HIP_TRY(hipGraphLaunch(plan->exec, s), "hipGraphLaunch");
    return PMLG_OK;
}

}  // namespace pmlg

// ---------------------------------------------------------------------------
// autotuner: times every valid configuration of one shape on the device
// ---------------------------------------------------------------------------
namespace pmlg {

static std::atomic<int> g_refine{-1};

bool refine_enabled()
{
    int v = g_refine.load();
    if (v < 0) {
        const char *e = getenv("PMLG_REFINE");
        v = (e && e[0] == '0') ? 0 : 1;
        g_refine.store(v);
    }
    return v != 0;
}

pmlg_status time_configs(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                         int reps, double *ms, int *best, unsigned mask)
{
    if (m <= 0 || k <= 0 || n <= 0 || lenA <= 0 || lenB <= 0 || reps <= 0 || !ms)
        return fail(PMLG_ERR_ARG, "autotune: bad shape");
    DeviceGuard g(ctx->device);
    hipStream_t s = ctx->stream;
    const int64_t lenC = lenA + lenB - 1;
    const int64_t N = (int64_t)1 << ceil_log2(lenC);
    // A plan runs each configuration's products inside launches shared with
    // many others, so a product is timed among `copies` identical ones (same
    // inputs, own outputs and evaluations) in one plan: enough copies that the
    // launches hold about 2^27 evaluation words, at most kTuneCopies.
    const double eval_words = (double)(m * k + k * n + m * n) * (double)N;
    const int copies = (int)std::min<double>(kTuneCopies, std::max(1.0, std::ceil(134217728.0 / eval_words)));
    const size_t in_words = (size_t)(m * k * lenA + k * n * lenB);
    const size_t out_words = (size_t)(m * n * lenC);
    // evaluation arena bound: 4 buffers of the largest entry count at the transform length
    const int64_t ent = std::max(m * k, std::max(k * n, m * n));
    const size_t ws_bytes = (size_t)(4 * ent * N) * 8 * copies + 4096 * 4 * copies + (1 << 20);
    std::lock_guard<std::mutex> lk(ctx->tune_mu);
    HIP_TRY(ctx->tune_io.ensure((in_words + out_words * copies) * 8 + 256), "hipMalloc(autotune io)");
    HIP_TRY(ctx->tune_ws.ensure(ws_bytes), "hipMalloc(autotune workspace)");
    uint64_t *A = static_cast<uint64_t *>(ctx->tune_io.ptr);
    uint64_t *B = A + m * k * lenA;
    uint64_t *C = B + k * n * lenB;
    hipLaunchKernelGGL(fill_random_kernel, dim3(grid_for((int64_t)in_words)), dim3(256), 0, s, A,
                       (int64_t)in_words, ctx->p, (uint64_t)(m * 1000003 + k * 1009 + n));
    HIP_TRY(hipGetLastError(), "fill_random");
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&e1), "hipEventCreate");
    double best_ms = 0;
    int best_c = -1;
    pmlg_status result = PMLG_OK;
    for (int c = 0; c < kNumConfigs && result == PMLG_OK; ++c) {
        ms[c] = -1.0;
        if (!(mask & (1u << c)) || !config_valid(ctx, m, k, n, lenA, lenB, c))
            continue;
        // direct convolution is only timed where it can compete
        if (kConfigs[c].school &&
            (double)m * n * k * lenC * std::min(lenA, lenB) > kSchoolTimeCap)
            continue;
        std::vector<pmlg_mul_desc> d((size_t)copies);
        for (int i = 0; i < copies; ++i) {
            d[(size_t)i] = pmlg_mul_desc{};
            d[(size_t)i].ctx = ctx;
            d[(size_t)i].A = A;
            d[(size_t)i].B = B;
            d[(size_t)i].C = C + (size_t)i * out_words;
            d[(size_t)i].m = m;
            d[(size_t)i].k = k;
            d[(size_t)i].n = n;
            d[(size_t)i].lenA = lenA;
            d[(size_t)i].lenB = lenB;
            d[(size_t)i].method = PMLG_METHOD_AUTO;
            d[(size_t)i].config = c;
        }
        pmlg_batch_plan plan = nullptr;
        pmlg_status st = build_plan(&plan, d.data(), copies, 0, ctx->tune_ws.ptr, ctx->tune_ws.bytes,
                                    nullptr);
        if (st != PMLG_OK) {
            result = st;
            break;
        }
        // warm-up run, then `reps` timed runs on the context's stream
        st = run_plan(plan, s);
        if (st == PMLG_OK) {
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < reps && st == PMLG_OK; ++r)
                st = run_plan(plan, s);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float t = 0;
            (void)hipEventElapsedTime(&t, e0, e1);
            ms[c] = (double)t / reps / copies;
            if (best_c < 0 || ms[c] < best_ms) {
                best_ms = ms[c];
                best_c = c;
            }
        }
        (void)hipStreamSynchronize(s);
        free_plan_memory(plan);
        delete plan;
        result = st;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (result != PMLG_OK)
        return result;
    if (best)
        *best = best_c;
    if (best_c >= 0 && mask == kAllConfigs)
        ctx->tuned[tune_key(ctx, m, k, n, lenA, lenB)] = best_c;
    return PMLG_OK;
}

}  // namespace pmlg

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------
extern "C" {

pmlg_status pmlg_batch_plan_create(pmlg_batch_plan *plan, const pmlg_mul_desc *descs,
                                   int64_t count, int flags)
{
    return build_plan(plan, descs, count, flags, nullptr, 0, nullptr);
}

pmlg_status pmlg_batch_plan_run(pmlg_batch_plan plan, void *stream)
{
    if (!plan)
        return fail(PMLG_ERR_ARG, "null plan");
    return run_plan(plan, stream ? (hipStream_t)stream : plan->stream);
}

pmlg_status pmlg_batch_plan_sync(pmlg_batch_plan plan)
{
    if (!plan)
        return fail(PMLG_ERR_ARG, "null plan");
    DeviceGuard g(plan->device);
    HIP_TRY(hipStreamSynchronize(plan->stream), "hipStreamSynchronize");
    return PMLG_OK;
}

pmlg_status pmlg_batch_plan_profile(pmlg_batch_plan plan, int reps, double *ms)
{
    if (!plan || reps <= 0 || !ms)
        return fail(PMLG_ERR_ARG, "batch_plan_profile: bad arguments");
    DeviceGuard g(plan->device);
    StageEvents se{};
    for (int i = 0; i < 4; ++i)
        HIP_TRY(hipEventCreate(&se.ev[i]), "hipEventCreate");
    struct Destroy {
        StageEvents &e;
        ~Destroy()
        {
            for (int i = 0; i < 4; ++i)
                (void)hipEventDestroy(e.ev[i]);
        }
    } d{se};
    double acc[3] = {0, 0, 0};
    for (int r = 0; r < reps; ++r) {
        pmlg_status st = launch_all(plan, plan->stream, &se);
        if (st != PMLG_OK)
            return st;
        HIP_TRY(hipEventSynchronize(se.ev[3]), "hipEventSynchronize");
        for (int i = 0; i < 3; ++i) {
            float t = 0;
            HIP_TRY(hipEventElapsedTime(&t, se.ev[i], se.ev[i + 1]), "hipEventElapsedTime");
            acc[i] += t;
        }
    }
    for (int i = 0; i < 3; ++i)
        ms[i] = acc[i] / reps;
    return PMLG_OK;
}

pmlg_status pmlg_batch_plan_time_stages(pmlg_batch_plan plan, int ring)
{
    if (!plan || ring < 0 || ring > 4096)
        return fail(PMLG_ERR_ARG, "batch_plan_time_stages: bad arguments");
    if (!(plan->flags & PMLG_BATCH_GRAPH))
        return fail(PMLG_ERR_ARG, "batch_plan_time_stages: the plan is not captured as a graph");
    DeviceGuard g(plan->device);
    (void)hipStreamSynchronize(plan->stream);
    free_ring(plan);
    if (plan->exec)
        (void)hipGraphExecDestroy(plan->exec);
    if (plan->graph)
        (void)hipGraphDestroy(plan->graph);
    plan->exec = nullptr;
    plan->graph = nullptr;
    plan->ring = ring;
    return PMLG_OK;
}

pmlg_status pmlg_batch_plan_stage_times(pmlg_batch_plan plan, double *ms, int64_t *replays)
{
    if (!plan || !ms)
        return fail(PMLG_ERR_ARG, "batch_plan_stage_times: bad arguments");
    DeviceGuard g(plan->device);
    double acc[3] = {0, 0, 0};
    int64_t n = 0;
    for (size_t r = 0; r < plan->rev.size(); ++r) {
        if (plan->replays[r] == 0)
            continue;
        HIP_TRY(hipEventSynchronize(plan->rev[r][3]), "hipEventSynchronize");
        for (int i = 0; i < 3; ++i) {
            float t = 0;
            HIP_TRY(hipEventElapsedTime(&t, plan->rev[r][i], plan->rev[r][i + 1]),
                    "hipEventElapsedTime");
            acc[i] += t;
        }
        ++n;
        plan->replays[r] = 0;
    }
    for (int i = 0; i < 3; ++i)
        ms[i] = n ? acc[i] / (double)n : 0.0;
    if (replays)
        *replays = n;
    return PMLG_OK;
}

pmlg_status pmlg_batch_plan_config(pmlg_batch_plan plan, int64_t i, int *config)
{
    if (!plan || !config || i < 0 || i >= (int64_t)plan->prods.size())
        return fail(PMLG_ERR_ARG, "batch_plan_config: bad arguments");
    *config = plan->prods[i].config;
    return PMLG_OK;
}

pmlg_status pmlg_batch_plan_stats(pmlg_batch_plan plan, int64_t *launches, uint64_t *arena_bytes)
{
    if (!plan)
        return fail(PMLG_ERR_ARG, "null plan");
    int64_t l = (int64_t)plan->pw.size() + (int64_t)plan->school.size() + (plan->nfill ? 1 : 0) +
                (int64_t)plan->fwd.size();
    for (const InvGroup &g : plan->inv)
        l += 1 + (g.mblocks > 0 ? 1 : 0);
    if (launches)
        *launches = l;
    if (arena_bytes)
        *arena_bytes = plan->arena_bytes;
    return PMLG_OK;
}

void pmlg_batch_plan_destroy(pmlg_batch_plan plan)
{
    if (!plan)
        return;
    free_plan_memory(plan);
    delete plan;
}

pmlg_status pmlg_polmat_mul_batch(const pmlg_mul_desc *descs, int64_t count, void *stream)
{
    pmlg_batch_plan plan = nullptr;
    pmlg_status st = build_plan(&plan, descs, count, 0, nullptr, 0, nullptr);
    if (st != PMLG_OK)
        return st;
    hipStream_t s = stream ? (hipStream_t)stream : plan->stream;
    st = run_plan(plan, s);
    if (st == PMLG_OK) {
        DeviceGuard g(plan->device);
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess)
            st = hip_fail(e, "hipStreamSynchronize");
    }
    pmlg_batch_plan_destroy(plan);
    return st;
}

pmlg_status pmlg_autotune(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA,
                          int64_t lenB, int reps, double *ms_per_config, int *best_config)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    double tmp[kNumConfigs];
    return time_configs(ctx, m, k, n, lenA, lenB, reps, ms_per_config ? ms_per_config : tmp,
                        best_config);
}

pmlg_status pmlg_fill_random(pmlg_ctx ctx, uint64_t *dst, int64_t n, uint64_t seed, void *stream)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    if (n < 0 || (n > 0 && !dst))
        return fail(PMLG_ERR_ARG, "fill_random: bad arguments");
    if (n == 0)
        return PMLG_OK;
    DeviceGuard g(ctx->device);
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    hipLaunchKernelGGL(fill_random_kernel, dim3(grid_for(n)), dim3(256), 0, s, dst, n, ctx->p, seed);
    HIP_TRY(hipGetLastError(), "fill_random");
    return PMLG_OK;
}

void pmlg_dispatch_set_refine(int on)
{
    g_refine.store(on ? 1 : 0);
}

int pmlg_num_configs(void)
{
    return kNumConfigs;
}

int pmlg_config_valid(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                      int config)
{
    return ctx && config_valid(ctx, m, k, n, lenA, lenB, config) ? 1 : 0;
}

int pmlg_choose_config(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                       int method)
{
    if (!ctx || m <= 0 || k <= 0 || n <= 0 || lenA <= 0 || lenB <= 0)
        return -1;
    return choose_config(ctx, m, k, n, lenA, lenB, method);
}

}  // extern "C"
