// Learned cost model of the batched dispatch: a small fp32 MLP
//   features (m/64, k/64, n/64, log2(product degree + 1)/16, bits(p)/64)
//   -> 64 -> 64 -> 64 (ReLU) -> one predicted log(milliseconds) per configuration,
// trained on timings the library records itself (pmlg_autotune) and run by
// the kernels below (forward, backward, Adam).  The dispatcher takes the
// valid configuration of least predicted time; the threshold rule stays the
// fallback.  The model only chooses among exact algorithms: whatever it
// picks, the product is the same.
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "dispatch.hpp"

namespace pmlg {
namespace {

constexpr int kLayers = 4;
constexpr int kDims[kLayers + 1] = {PMLG_MLP_IN, PMLG_MLP_HIDDEN, PMLG_MLP_HIDDEN, PMLG_MLP_HIDDEN,
                                    PMLG_MLP_OUT};
constexpr int kMaxWidth = 64;
constexpr int kThreads = 256;

constexpr int param_count()
{
    int n = 0;
    for (int l = 0; l < kLayers; ++l)
        n += kDims[l] * kDims[l + 1] + kDims[l + 1];
    return n;
}
constexpr int kParams = param_count();
static_assert(kParams == PMLG_MLP_PARAMS, "parameter count");

// offsets of W_l (in x out, row-major) and b_l in the flat parameter vector
constexpr int w_off(int l)
{
    int n = 0;
    for (int i = 0; i < l; ++i)
        n += kDims[i] * kDims[i + 1] + kDims[i + 1];
    return n;
}
constexpr int b_off(int l)
{
    return w_off(l) + kDims[l] * kDims[l + 1];
}

// Z = X W + b, A = relu(Z) (or Z for the output layer); one thread per (row, unit)
__global__ void __launch_bounds__(kThreads)
linear_fwd(const float *__restrict__ X, int in, const float *__restrict__ W,
           const float *__restrict__ b, float *__restrict__ Z, float *__restrict__ A, int out,
           int rows, int relu)
{
    const int idx = blockIdx.x * kThreads + threadIdx.x;
    if (idx >= rows * out)
        return;
    const int r = idx / out, o = idx - r * out;
    float z = b[o];
    for (int i = 0; i < in; ++i)
        z = fmaf(X[r * in + i], W[i * out + o], z);
    Z[idx] = z;
    A[idx] = relu ? fmaxf(z, 0.0f) : z;
}

// dP = 2 M (P - Y) / count; per-row partial losses (summed by loss_sum)
__global__ void __launch_bounds__(kThreads)
loss_grad(const float *__restrict__ P, const float *__restrict__ Y, const float *__restrict__ M,
          float *__restrict__ dP, float *__restrict__ row_loss, int rows, int K, float inv_count)
{
    const int r = blockIdx.x * kThreads + threadIdx.x;
    if (r >= rows)
        return;
    float acc = 0.0f;
    for (int c = 0; c < K; ++c) {
        const float d = (P[r * K + c] - Y[r * K + c]) * M[r * K + c];
        acc = fmaf(d, d, acc);
        dP[r * K + c] = 2.0f * d * inv_count;
    }
    row_loss[r] = acc * inv_count;
}

// one workgroup: loss = sum of the per-row partials (fixed order)
__global__ void __launch_bounds__(kThreads) loss_sum(const float *__restrict__ row_loss, int rows,
                                                     float *__restrict__ loss)
{
    __shared__ float part[kThreads];
    float a = 0.0f;
    for (int r = threadIdx.x; r < rows; r += kThreads)
        a += row_loss[r];
    part[threadIdx.x] = a;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *loss = part[0];
}

// dX = dZ W^T, then masked by relu'(Zprev) when the previous layer has a ReLU
__global__ void __launch_bounds__(kThreads)
linear_bwd_dx(const float *__restrict__ dZ, const float *__restrict__ W, int in, int out,
              const float *__restrict__ Zprev, float *__restrict__ dX, int rows)
{
    const int idx = blockIdx.x * kThreads + threadIdx.x;
    if (idx >= rows * in)
        return;
    const int r = idx / in, i = idx - r * in;
    float g = 0.0f;
    for (int o = 0; o < out; ++o)
        g = fmaf(dZ[r * out + o], W[i * out + o], g);
    if (Zprev && Zprev[idx] <= 0.0f)
        g = 0.0f;
    dX[idx] = g;
}

// dW = X^T dZ, db = sum_r dZ; one thread per (i, o), thread row i == in for db
__global__ void __launch_bounds__(kThreads)
linear_bwd_dw(const float *__restrict__ X, const float *__restrict__ dZ, int in, int out, int rows,
              float *__restrict__ dW, float *__restrict__ db)
{
    const int idx = blockIdx.x * kThreads + threadIdx.x;
    if (idx >= (in + 1) * out)
        return;
    const int i = idx / out, o = idx - i * out;
    float g = 0.0f;
    if (i < in) {
        for (int r = 0; r < rows; ++r)
            g = fmaf(X[r * in + i], dZ[r * out + o], g);
        dW[i * out + o] = g;
    } else {
        for (int r = 0; r < rows; ++r)
            g += dZ[r * out + o];
        db[o] = g;
    }
}

__global__ void __launch_bounds__(kThreads)
adam_step(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
          float *__restrict__ m2, int n, float lr, float b1, float b2, float eps, float bc1,
          float bc2)
{
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n)
        return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float v = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    w[i] -= lr * (a / bc1) / (sqrtf(v / bc2) + eps);
}

// rows of the training set picked by idx (a shuffled batch)
__global__ void __launch_bounds__(kThreads)
gather_rows(const float *__restrict__ src, const int *__restrict__ idx, int rows, int width,
            float *__restrict__ dst)
{
    const int t = blockIdx.x * kThreads + threadIdx.x;
    if (t >= rows * width)
        return;
    const int r = t / width, c = t - r * width;
    dst[t] = src[(size_t)idx[r] * width + c];
}

unsigned blocks_for(int n)
{
    return (unsigned)((n + kThreads - 1) / kThreads);
}

}  // namespace
}  // namespace pmlg

using namespace pmlg;

struct pmlg_mlp_s {
    int device = 0;
    hipStream_t stream = nullptr;
    float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
    int64_t adam_t = 0;
    // activations for up to `cap` rows: Z_l and A_l per layer, input copy, dP
    int cap = 0;
    float *act = nullptr;
    float *loss = nullptr;
    std::mutex mu;
};

namespace {

struct Acts {
    float *X, *Y, *M, *Z[kLayers], *A[kLayers], *dA, *dB, *rl;
};

// carve the activation block for `rows` rows (cap >= rows)
Acts carve(pmlg_mlp m)
{
    Acts a{};
    float *p = m->act;
    const int R = m->cap;
    a.X = p;
    p += R * kDims[0];
    a.Y = p;
    p += R * PMLG_MLP_OUT;
    a.M = p;
    p += R * PMLG_MLP_OUT;
    for (int l = 0; l < kLayers; ++l) {
        a.Z[l] = p;
        p += R * kDims[l + 1];
        a.A[l] = p;
        p += R * kDims[l + 1];
    }
    a.dA = p;
    p += R * kMaxWidth;
    a.dB = p;
    p += R * kMaxWidth;
    a.rl = p;
    return a;
}

size_t act_floats(int R)
{
    size_t n = (size_t)R * (kDims[0] + 2 * PMLG_MLP_OUT + 2 * kMaxWidth + 1);
    for (int l = 0; l < kLayers; ++l)
        n += 2 * (size_t)R * kDims[l + 1];
    return n;
}

pmlg_status ensure_cap(pmlg_mlp m, int rows)
{
    if (rows <= m->cap)
        return PMLG_OK;
    if (m->act) {
        (void)hipStreamSynchronize(m->stream);
        (void)hipFree(m->act);
        m->act = nullptr;
    }
    int R = std::max(rows, 256);
    HIP_TRY(hipMalloc(&m->act, act_floats(R) * sizeof(float)), "hipMalloc(mlp activations)");
    m->cap = R;
    return PMLG_OK;
}

pmlg_status forward(pmlg_mlp m, const Acts &a, int rows)
{
    const float *x = a.X;
    for (int l = 0; l < kLayers; ++l) {
        const int in = kDims[l], out = kDims[l + 1];
        hipLaunchKernelGGL(linear_fwd, dim3(blocks_for(rows * out)), dim3(kThreads), 0, m->stream, x,
                           in, m->params + w_off(l), m->params + b_off(l), a.Z[l], a.A[l], out, rows,
                           l + 1 < kLayers ? 1 : 0);
        x = a.A[l];
    }
    HIP_TRY(hipGetLastError(), "mlp forward");
    return PMLG_OK;
}

// forward + loss + backward into m->grads (a.X, a.Y, a.M filled)
pmlg_status loss_and_grads(pmlg_mlp m, const Acts &a, int rows, float inv_count)
{
    pmlg_status st = forward(m, a, rows);
    if (st != PMLG_OK)
        return st;
    const int K = PMLG_MLP_OUT;
    hipLaunchKernelGGL(loss_grad, dim3(blocks_for(rows)), dim3(kThreads), 0, m->stream,
                       a.A[kLayers - 1], a.Y, a.M, a.dA, a.rl, rows, K, inv_count);
    hipLaunchKernelGGL(loss_sum, dim3(1), dim3(kThreads), 0, m->stream, a.rl, rows, m->loss);
    // dA holds dZ of the output layer (no activation there)
    float *dz = a.dA, *dnext = a.dB;
    for (int l = kLayers - 1; l >= 0; --l) {
        const int in = kDims[l], out = kDims[l + 1];
        const float *xin = l > 0 ? a.A[l - 1] : a.X;
        hipLaunchKernelGGL(linear_bwd_dw, dim3(blocks_for((in + 1) * out)), dim3(kThreads), 0,
                           m->stream, xin, dz, in, out, rows, m->grads + w_off(l),
                           m->grads + b_off(l));
        if (l > 0) {
            hipLaunchKernelGGL(linear_bwd_dx, dim3(blocks_for(rows * in)), dim3(kThreads), 0,
                               m->stream, dz, m->params + w_off(l), in, out, a.Z[l - 1], dnext,
                               rows);
            std::swap(dz, dnext);
        }
    }
    HIP_TRY(hipGetLastError(), "mlp backward");
    return PMLG_OK;
}

std::mutex g_model_mu;
pmlg_mlp g_model = nullptr;

}  // namespace

namespace pmlg {

void mlp_features(int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB, uint64_t p, float *x)
{
    const int64_t lenC = lenA + lenB - 1;
    x[0] = (float)m / 64.0f;
    x[1] = (float)k / 64.0f;
    x[2] = (float)n / 64.0f;
    x[3] = (float)std::log2((double)std::max<int64_t>(lenC, 1)) / 16.0f;
    x[4] = (float)(64 - __builtin_clzll(p)) / 64.0f;
}

int learned_config(pmlg_ctx ctx, int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB,
                   float *pred)
{
    std::lock_guard<std::mutex> lk(g_model_mu);
    if (!g_model)
        return -1;
    float x[PMLG_MLP_IN], y[PMLG_MLP_OUT];
    mlp_features(m, k, n, lenA, lenB, ctx->p, x);
    if (pmlg_mlp_forward(g_model, x, 1, y) != PMLG_OK)
        return -1;
    int best = -1;
    for (int c = 0; c < PMLG_MLP_OUT && c < kNumConfigs; ++c) {
        const bool ok = config_valid(ctx, m, k, n, lenA, lenB, c);
        if (pred)
            pred[c] = ok ? y[c] : INFINITY;
        if (ok && (best < 0 || y[c] < y[best]))
            best = c;
    }
    return best;
}

}  // namespace pmlg

extern "C" {

pmlg_status pmlg_mlp_create(pmlg_mlp *out, int device, uint64_t seed)
{
    if (!out)
        return fail(PMLG_ERR_ARG, "null output pointer");
    *out = nullptr;
    pmlg_mlp m = new (std::nothrow) pmlg_mlp_s();
    if (!m)
        return fail(PMLG_ERR_ALLOC, "out of host memory");
    m->device = device;
    DeviceGuard g(device);
    auto bail = [&](hipError_t e, const char *w) {
        pmlg_mlp_destroy(m);
        return hip_fail(e, w);
    };
    hipError_t e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking);
    if (e != hipSuccess)
        return bail(e, "hipStreamCreate");
    float *mem = nullptr;
    e = hipMalloc(&mem, (4 * (size_t)kParams + 64) * sizeof(float));
    if (e != hipSuccess)
        return bail(e, "hipMalloc(mlp)");
    m->params = mem;
    m->grads = mem + kParams;
    m->m1 = mem + 2 * kParams;
    m->m2 = mem + 3 * kParams;
    m->loss = mem + 4 * kParams;
    // He-uniform weights, zero biases, zero Adam moments
    std::vector<float> h(4 * (size_t)kParams, 0.0f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < kLayers; ++l) {
        const float lim = std::sqrt(6.0f / (float)kDims[l]);
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int i = 0; i < kDims[l] * kDims[l + 1]; ++i)
            h[w_off(l) + i] = U(rng);
    }
    e = hipMemcpy(mem, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess)
        return bail(e, "hipMemcpy(mlp)");
    *out = m;
    return PMLG_OK;
}

void pmlg_mlp_destroy(pmlg_mlp m)
{
    if (!m)
        return;
    {
        std::lock_guard<std::mutex> lk(g_model_mu);
        if (g_model == m)
            g_model = nullptr;
    }
    DeviceGuard g(m->device);
    if (m->stream)
        (void)hipStreamSynchronize(m->stream);
    if (m->params)
        (void)hipFree(m->params);
    if (m->act)
        (void)hipFree(m->act);
    if (m->stream)
        (void)hipStreamDestroy(m->stream);
    delete m;
}

pmlg_status pmlg_mlp_get_params(pmlg_mlp m, float *params)
{
    if (!m || !params)
        return fail(PMLG_ERR_ARG, "mlp_get_params: bad arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    HIP_TRY(hipStreamSynchronize(m->stream), "sync");
    HIP_TRY(hipMemcpy(params, m->params, kParams * sizeof(float), hipMemcpyDeviceToHost), "D2H");
    return PMLG_OK;
}

pmlg_status pmlg_mlp_set_params(pmlg_mlp m, const float *params)
{
    if (!m || !params)
        return fail(PMLG_ERR_ARG, "mlp_set_params: bad arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    HIP_TRY(hipStreamSynchronize(m->stream), "sync");
    HIP_TRY(hipMemcpy(m->params, params, kParams * sizeof(float), hipMemcpyHostToDevice), "H2D");
    // a new parameter set restarts the optimiser state
    std::vector<float> z(2 * (size_t)kParams, 0.0f);
    HIP_TRY(hipMemcpy(m->m1, z.data(), z.size() * sizeof(float), hipMemcpyHostToDevice), "H2D");
    m->adam_t = 0;
    return PMLG_OK;
}

pmlg_status pmlg_mlp_forward(pmlg_mlp m, const float *X, int64_t rows, float *out)
{
    if (!m || rows < 0 || (rows > 0 && (!X || !out)) || rows > (1 << 24))
        return fail(PMLG_ERR_ARG, "mlp_forward: bad arguments");
    if (rows == 0)
        return PMLG_OK;
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    pmlg_status st = ensure_cap(m, (int)rows);
    if (st != PMLG_OK)
        return st;
    Acts a = carve(m);
    HIP_TRY(hipMemcpyAsync(a.X, X, rows * PMLG_MLP_IN * sizeof(float), hipMemcpyHostToDevice,
                           m->stream),
            "H2D");
    if ((st = forward(m, a, (int)rows)) != PMLG_OK)
        return st;
    HIP_TRY(hipMemcpyAsync(out, a.A[kLayers - 1], rows * PMLG_MLP_OUT * sizeof(float),
                           hipMemcpyDeviceToHost, m->stream),
            "D2H");
    HIP_TRY(hipStreamSynchronize(m->stream), "sync");
    return PMLG_OK;
}

pmlg_status pmlg_mlp_loss_grad(pmlg_mlp m, const float *X, const float *Y, const float *mask,
                               int64_t rows, float *loss, float *grads)
{
    if (!m || rows <= 0 || !X || !Y || !mask || rows > (1 << 20))
        return fail(PMLG_ERR_ARG, "mlp_loss_grad: bad arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    pmlg_status st = ensure_cap(m, (int)rows);
    if (st != PMLG_OK)
        return st;
    Acts a = carve(m);
    double cnt = 0;
    for (int64_t i = 0; i < rows * PMLG_MLP_OUT; ++i)
        cnt += mask[i];
    const float inv = cnt > 0 ? (float)(1.0 / cnt) : 0.0f;
    hipStream_t s = m->stream;
    HIP_TRY(hipMemcpyAsync(a.X, X, rows * PMLG_MLP_IN * sizeof(float), hipMemcpyHostToDevice, s), "H2D");
    HIP_TRY(hipMemcpyAsync(a.Y, Y, rows * PMLG_MLP_OUT * sizeof(float), hipMemcpyHostToDevice, s), "H2D");
    HIP_TRY(hipMemcpyAsync(a.M, mask, rows * PMLG_MLP_OUT * sizeof(float), hipMemcpyHostToDevice, s), "H2D");
    if ((st = loss_and_grads(m, a, (int)rows, inv)) != PMLG_OK)
        return st;
    if (loss)
        HIP_TRY(hipMemcpyAsync(loss, m->loss, sizeof(float), hipMemcpyDeviceToHost, s), "D2H");
    if (grads)
        HIP_TRY(hipMemcpyAsync(grads, m->grads, kParams * sizeof(float), hipMemcpyDeviceToHost, s), "D2H");
    HIP_TRY(hipStreamSynchronize(s), "sync");
    return PMLG_OK;
}

pmlg_status pmlg_mlp_train(pmlg_mlp m, const float *X, const float *Y, const float *mask,
                           int64_t rows, int epochs, int batch, float lr, uint64_t seed,
                           float *final_loss)
{
    if (!m || rows <= 0 || !X || !Y || !mask || epochs < 0 || batch <= 0 || batch > 4096 ||
        rows > (1 << 22))
        return fail(PMLG_ERR_ARG, "mlp_train: bad arguments");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(m->device);
    hipStream_t s = m->stream;
    // training set resident on the device
    const size_t xs = rows * PMLG_MLP_IN, ys = rows * PMLG_MLP_OUT;
    float *dset = nullptr;
    int *didx = nullptr;
    HIP_TRY(hipMalloc(&dset, (xs + 2 * ys) * sizeof(float)), "hipMalloc(training set)");
    struct Free {
        void *a, *b;
        ~Free()
        {
            if (a)
                (void)hipFree(a);
            if (b)
                (void)hipFree(b);
        }
    } fr{dset, nullptr};
    HIP_TRY(hipMalloc(&didx, rows * sizeof(int)), "hipMalloc(indices)");
    fr.b = didx;
    HIP_TRY(hipMemcpy(dset, X, xs * sizeof(float), hipMemcpyHostToDevice), "H2D");
    HIP_TRY(hipMemcpy(dset + xs, Y, ys * sizeof(float), hipMemcpyHostToDevice), "H2D");
    HIP_TRY(hipMemcpy(dset + xs + ys, mask, ys * sizeof(float), hipMemcpyHostToDevice), "H2D");
    pmlg_status st = ensure_cap(m, batch);
    if (st != PMLG_OK)
        return st;
    Acts a = carve(m);
    std::vector<int> idx(rows);
    for (int64_t i = 0; i < rows; ++i)
        idx[i] = (int)i;
    std::mt19937_64 rng(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float last = 0.0f;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(idx.begin(), idx.end(), rng);
        HIP_TRY(hipMemcpyAsync(didx, idx.data(), rows * sizeof(int), hipMemcpyHostToDevice, s), "H2D");
        // the host copy of idx is reshuffled only after this epoch's copy is done
        for (int64_t r0 = 0; r0 + batch <= rows || (r0 == 0 && rows < batch); r0 += batch) {
            const int nb = (int)std::min<int64_t>(batch, rows - r0);
            // mask count of this batch, for the mean
            double cnt = 0;
            for (int i = 0; i < nb; ++i)
                for (int c = 0; c < PMLG_MLP_OUT; ++c)
                    cnt += mask[(size_t)idx[r0 + i] * PMLG_MLP_OUT + c];
            const float inv = cnt > 0 ? (float)(1.0 / cnt) : 0.0f;
            hipLaunchKernelGGL(gather_rows, dim3(blocks_for(nb * PMLG_MLP_IN)), dim3(kThreads), 0, s,
                               dset, didx + r0, nb, PMLG_MLP_IN, a.X);
            hipLaunchKernelGGL(gather_rows, dim3(blocks_for(nb * PMLG_MLP_OUT)), dim3(kThreads), 0, s,
                               dset + xs, didx + r0, nb, PMLG_MLP_OUT, a.Y);
            hipLaunchKernelGGL(gather_rows, dim3(blocks_for(nb * PMLG_MLP_OUT)), dim3(kThreads), 0, s,
                               dset + xs + ys, didx + r0, nb, PMLG_MLP_OUT, a.M);
            if ((st = loss_and_grads(m, a, nb, inv)) != PMLG_OK)
                return st;
            ++m->adam_t;
            const float bc1 = 1.0f - std::pow(b1, (float)m->adam_t);
            const float bc2 = 1.0f - std::pow(b2, (float)m->adam_t);
            hipLaunchKernelGGL(adam_step, dim3(blocks_for(kParams)), dim3(kThreads), 0, s, m->params,
                               m->grads, m->m1, m->m2, kParams, lr, b1, b2, eps, bc1, bc2);
            HIP_TRY(hipGetLastError(), "adam");
        }
        HIP_TRY(hipStreamSynchronize(s), "sync");
    }
    // loss of the last parameters over the whole set
    if (final_loss) {
        double sum = 0, cnt = 0;
        std::vector<float> out(PMLG_MLP_OUT * 4096);
        for (int64_t r0 = 0; r0 < rows; r0 += 4096) {
            const int nb = (int)std::min<int64_t>(4096, rows - r0);
            if ((st = ensure_cap(m, nb)) != PMLG_OK)
                return st;
            a = carve(m);
            HIP_TRY(hipMemcpyAsync(a.X, X + r0 * PMLG_MLP_IN, nb * PMLG_MLP_IN * sizeof(float),
                                   hipMemcpyHostToDevice, s),
                    "H2D");
            if ((st = forward(m, a, nb)) != PMLG_OK)
                return st;
            HIP_TRY(hipMemcpyAsync(out.data(), a.A[kLayers - 1], nb * PMLG_MLP_OUT * sizeof(float),
                                   hipMemcpyDeviceToHost, s),
                    "D2H");
            HIP_TRY(hipStreamSynchronize(s), "sync");
            for (int i = 0; i < nb * PMLG_MLP_OUT; ++i) {
                const double mk = mask[r0 * PMLG_MLP_OUT + i];
                const double d = (out[i] - Y[r0 * PMLG_MLP_OUT + i]) * mk;
                sum += d * d;
                cnt += mk;
            }
        }
        last = cnt > 0 ? (float)(sum / cnt) : 0.0f;
        *final_loss = last;
    }
    return PMLG_OK;
}

pmlg_status pmlg_dispatch_set_model(pmlg_mlp m)
{
    std::lock_guard<std::mutex> lk(g_model_mu);
    g_model = m;
    return PMLG_OK;
}

void pmlg_mlp_features(int64_t m, int64_t k, int64_t n, int64_t lenA, int64_t lenB, uint64_t p,
                       float *x)
{
    mlp_features(m, k, n, lenA, lenB, p, x);
}

}  // extern "C"
