// Minimal left kernel bases of polynomial matrices over Z/pZ (shifted ordered
// weak Popov form), device-resident.  Two algorithms, as in
// flint-extras/src/nmod_poly_mat_extra/kernel.c:
//   - Zhou-Labahn-Storjohann (nmod_poly_mat_kernel_zls_approx, kernel.c:232-478):
//     column splitting while 2n > m, otherwise one approximant basis at an
//     order chosen from the shift, a degree test that picks the rows already
//     in the kernel, and a recursion on the residual of the other rows;
//   - one approximant basis at an order large enough for the whole kernel
//     (nmod_poly_mat_kernel_via_approx, kernel.c:148-228).
// Every approximant basis, product and middle product runs on the device (the
// PM-Basis and product paths of this library) and the matrices stay in HBM;
// only entry degrees, shifts and pivot indices come back to the host, where
// the recursion is steered.
#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "ctx.hpp"

namespace pmlg {

namespace {

// deg[e] = degree of polynomial e (coefficients at d + e*cap, length len), -1
// for zero.  One wave per entry reads 64 coefficients at a time from the top
// and takes the highest nonzero one from a ballot.
__global__ void __launch_bounds__(256) entry_degree_kernel(const uint64_t *__restrict__ d,
                                                           int64_t cap, int64_t len,
                                                           int64_t entries, int *__restrict__ deg)
{
    const int lane = threadIdx.x & 63;
    const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (e >= entries)
        return;
    const uint64_t *x = d + e * cap;
    int result = -1;
    for (int64_t top = len - 1; top >= 0; top -= 64) {
        const int64_t i = top - lane;
        const unsigned long long mask = __ballot(i >= 0 && x[i] != 0);
        if (mask) {
            result = (int)(top - (__ffsll(mask) - 1));
            break;
        }
    }
    if (lane == 0)
        deg[e] = result;
}

constexpr int kMaxRows = 256;  // rows of F (the approximant bases' limit)
constexpr int kMaxApproxCols = 128;  // columns of one approximant basis

// Row moves of one launch: destination row dst[t] <- source row src[t].
struct RowMap {
    int dst[kMaxRows];
    int src[kMaxRows];
};

// dst (dc columns, capacity dcap): row map.dst[t], columns [0, ncols) receive
// src (sc columns, capacity scap) row map.src[t], columns [col0, col0+ncols);
// coefficients [0, slen) are copied and [slen, dlen) zeroed.
__global__ void __launch_bounds__(256) gather_rows_kernel(uint64_t *__restrict__ dst, int64_t dcap,
                                                          int64_t dc, int64_t dlen,
                                                          const uint64_t *__restrict__ src,
                                                          int64_t scap, int64_t sc, int64_t slen,
                                                          int64_t col0, int64_t ncols, RowMap map,
                                                          int64_t rows)
{
    const int64_t per_row = ncols * dlen;
    const int64_t items = rows * per_row;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < items;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = idx / per_row;
        const int64_t rem = idx - t * per_row;
        const int64_t j = rem / dlen, k = rem - j * dlen;
        const int64_t dr = map.dst[t], sr = map.src[t];
        dst[(dr * dc + j) * dcap + k] = k < slen ? src[(sr * sc + col0 + j) * scap + k] : 0;
    }
}

// m x m identity with capacity cap (coefficient 0 only)
__global__ void __launch_bounds__(256) identity_kernel(uint64_t *__restrict__ dst, int64_t cap,
                                                       int64_t m)
{
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < m * m;
         idx += (int64_t)gridDim.x * blockDim.x)
        dst[idx * cap] = (idx / m == idx % m) ? 1 : 0;
}

unsigned grid_of(int64_t items)
{
    int64_t b = (items + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 8192));
}

// r x c polynomial matrix in device memory, entry (i, j) coefficient k at
// d[(i*c + j)*cap + k]; only coefficients below len are meaningful.
struct DMat {
    std::shared_ptr<uint64_t> mem;
    uint64_t *d = nullptr;
    int64_t r = 0, c = 0, cap = 0, len = 0;
};

DMat rows_of(const DMat &M, int64_t r0, int64_t nr)
{
    DMat w = M;
    w.d = M.d + r0 * M.c * M.cap;
    w.r = nr;
    return w;
}

struct Kx {
    pmlg_ctx ctx;
    hipStream_t s;
    int method;
    DevBuf ibuf;  // entry degrees
    ~Kx() { ibuf.release(); }
};

pmlg_status alloc(DMat &M, int64_t r, int64_t c, int64_t cap)
{
    void *p = nullptr;
    const size_t words = (size_t)std::max<int64_t>(r * c * cap, 1);
    HIP_TRY(hipMalloc(&p, words * 8), "hipMalloc(kernel basis)");
    M.mem = std::shared_ptr<uint64_t>(static_cast<uint64_t *>(p), [](uint64_t *q) { (void)hipFree(q); });
    M.d = static_cast<uint64_t *>(p);
    M.r = r;
    M.c = c;
    M.cap = cap;
    M.len = 0;
    return PMLG_OK;
}

// degrees of every entry (row-major), synchronous
pmlg_status entry_degrees(Kx &X, const DMat &M, std::vector<int> &deg)
{
    const int64_t entries = M.r * M.c;
    deg.assign((size_t)entries, -1);
    if (entries == 0)
        return PMLG_OK;
    HIP_TRY(X.ibuf.ensure((size_t)entries * sizeof(int)), "hipMalloc(degrees)");
    int *dd = static_cast<int *>(X.ibuf.ptr);
    hipLaunchKernelGGL(entry_degree_kernel, dim3((unsigned)((entries + 3) / 4)), dim3(256), 0, X.s,
                       M.d, M.cap, M.len, entries, dd);
    HIP_TRY(hipGetLastError(), "entry_degree_kernel");
    HIP_TRY(hipMemcpyAsync(deg.data(), dd, (size_t)entries * sizeof(int), hipMemcpyDeviceToHost, X.s),
            "D2H(degrees)");
    HIP_TRY(hipStreamSynchronize(X.s), "sync");
    return PMLG_OK;
}

int max_of(const std::vector<int> &v)
{
    int mx = -1;
    for (int x : v)
        mx = std::max(mx, x);
    return mx;
}

// sets M.len to its normalised length (at least 1)
pmlg_status trim(Kx &X, DMat &M, std::vector<int> *out = nullptr)
{
    std::vector<int> deg;
    pmlg_status st = entry_degrees(X, M, deg);
    if (st != PMLG_OK)
        return st;
    M.len = std::max(1, max_of(deg) + 1);
    if (out)
        out->swap(deg);
    return PMLG_OK;
}

// dst rows dr[t] <- src rows sr[t], columns [col0, col0 + ncols)
pmlg_status gather(Kx &X, DMat &dst, int64_t dlen, const DMat &src, int64_t col0, int64_t ncols,
                   const std::vector<int> &dr, const std::vector<int> &sr)
{
    const int64_t rows = (int64_t)dr.size();
    if (rows == 0 || ncols == 0 || dlen == 0)
        return PMLG_OK;
    if (rows > kMaxRows || ncols > dst.c || col0 + ncols > src.c || dlen > dst.cap)
        return fail(PMLG_ERR_ARG, "kernel basis: internal row copy out of range");
    RowMap map{};
    for (int64_t t = 0; t < rows; ++t) {
        if (dr[t] < 0 || dr[t] >= dst.r || sr[t] < 0 || sr[t] >= src.r)
            return fail(PMLG_ERR_ARG, "kernel basis: internal row index out of range");
        map.dst[t] = dr[t];
        map.src[t] = sr[t];
    }
    const int64_t slen = std::min(src.len, dlen);
    hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_of(rows * ncols * dlen)), dim3(256), 0, X.s,
                       dst.d, dst.cap, dst.c, dlen, src.d, src.cap, src.c, slen, col0, ncols, map,
                       rows);
    HIP_TRY(hipGetLastError(), "gather_rows_kernel");
    return PMLG_OK;
}

std::vector<int> iota_vec(int64_t n, int start = 0)
{
    std::vector<int> v((size_t)n);
    for (int64_t i = 0; i < n; ++i)
        v[(size_t)i] = (int)(start + i);
    return v;
}

// columns [c0, c0 + nc) of F as a new matrix
pmlg_status columns(Kx &X, const DMat &F, int64_t c0, int64_t nc, DMat &out)
{
    pmlg_status st = alloc(out, F.r, nc, F.len);
    if (st != PMLG_OK)
        return st;
    out.len = F.len;
    std::vector<int> rows = iota_vec(F.r);
    return gather(X, out, F.len, F, c0, nc, rows, rows);
}

pmlg_status identity(Kx &X, int64_t m, DMat &K)
{
    pmlg_status st = alloc(K, m, m, 1);
    if (st != PMLG_OK)
        return st;
    K.len = 1;
    hipLaunchKernelGGL(identity_kernel, dim3(grid_of(m * m)), dim3(256), 0, X.s, K.d, (int64_t)1, m);
    HIP_TRY(hipGetLastError(), "identity_kernel");
    return PMLG_OK;
}

// C = A * B, normalised length
pmlg_status product(Kx &X, const DMat &A, const DMat &B, DMat &C)
{
    const int64_t lc = A.len + B.len - 1;
    pmlg_status st = alloc(C, A.r, B.c, lc);
    if (st != PMLG_OK)
        return st;
    C.len = lc;
    st = mulmid_dev_strided(X.ctx, C.d, C.cap, A.d, A.cap, B.d, B.cap, A.r, A.c, B.c, A.len, B.len, 0,
                            lc, X.method, X.s);
    if (st != PMLG_OK)
        return st;
    return trim(X, C);
}

#define KTRY(expr)                        \
    do {                                  \
        pmlg_status _st = (expr);         \
        if (_st != PMLG_OK)               \
            return _st;                   \
    } while (0)

// kernel.c:232-478.  F is m x n with F.len >= 1; shift (in/out) and pivind
// (out) have m entries.  K receives at least *nul rows, the first *nul of
// which form the basis.
pmlg_status zls(Kx &X, const DMat &Fin, int64_t *shift, int64_t *pivind, int64_t *nul, DMat &K)
{
    const int64_t m = Fin.r, n = Fin.c;
    std::vector<int> deg;
    KTRY(entry_degrees(X, Fin, deg));
    const int dmax = max_of(deg);
    DMat F = Fin;
    F.len = std::max(1, dmax + 1);
    if (m == 0 || n == 0 || ((m == 1 || n == 1) && dmax < 0)) {
        KTRY(identity(X, m, K));
        for (int64_t i = 0; i < m; ++i)
            pivind[i] = i;
        *nul = m;
        return PMLG_OK;
    }
    if (m == 1) {
        *nul = 0;
        return PMLG_OK;
    }

    if (2 * n > m) {
        // F = [F1 | F2], F1 with n - n/2 columns (kernel.c:252-313)
        const int64_t n1 = n - n / 2;
        DMat F1, K1;
        KTRY(columns(X, F, 0, n1, F1));
        int64_t nul1 = 0;
        KTRY(zls(X, F1, shift, pivind, &nul1, K1));
        F1 = DMat();
        if (nul1 == 0) {
            *nul = 0;
            return PMLG_OK;
        }
        const DMat ker1 = rows_of(K1, 0, nul1);
        DMat F2, R;
        KTRY(columns(X, F, n1, n / 2, F2));
        KTRY(product(X, ker1, F2, R));
        F2 = DMat();
        std::vector<int64_t> piv2((size_t)nul1);
        int64_t nul2 = 0;
        DMat K2;
        KTRY(zls(X, R, shift, piv2.data(), &nul2, K2));
        R = DMat();
        if (nul2 == 0) {
            *nul = 0;
            return PMLG_OK;
        }
        KTRY(product(X, rows_of(K2, 0, nul2), ker1, K));
        for (int64_t i = 0; i < nul2; ++i)
            pivind[i] = pivind[piv2[(size_t)i]];
        *nul = nul2;
        return PMLG_OK;
    }

    // 1 <= n <= m/2 (kernel.c:315-477)
    std::vector<int64_t> buf((size_t)m);
    int64_t nnz = 0;
    for (int64_t i = 0; i < m; ++i) {
        int rm = -1;
        for (int64_t j = 0; j < n; ++j)
            rm = std::max(rm, deg[(size_t)(i * n + j)]);
        buf[(size_t)i] = rm >= 0 ? rm : 0;
        nnz += rm >= 0;
    }
    if (nnz == 0) {
        KTRY(identity(X, m, K));
        for (int64_t i = 0; i < m; ++i)
            pivind[i] = i;
        *nul = m;
        return PMLG_OK;
    }
    int64_t maxdeg = buf[0], diff_shift = shift[0] - buf[0];
    for (int64_t i = 1; i < m; ++i) {
        maxdeg = std::max(maxdeg, buf[(size_t)i]);
        diff_shift = std::min(diff_shift, shift[i] - buf[(size_t)i]);
    }
    // order = 1 + max(maxdeg, 1 + floor((sum(shift) - m diff_shift - 1) / (m - n)))
    // with C's truncating division, as kernel.c:349-353
    int64_t order = -m * diff_shift - 1;
    for (int64_t i = 0; i < m; ++i)
        order += shift[i];
    order = 1 + order / (m - n);
    order = 1 + std::max(maxdeg, order);

    DMat P;
    KTRY(alloc(P, m, m, order + 1));
    int64_t lenP = 0;
    KTRY(basis_common(X.ctx, P.d, P.cap, &lenP, shift, F.d, F.cap, m, n, F.len, order, X.method,
                      false, X.s, nullptr));
    P.len = std::max<int64_t>(lenP, 1);
    std::vector<int> pdeg;
    KTRY(trim(X, P, &pdeg));

    // rows already in the kernel, then the others (kernel.c:361-386)
    int64_t nk = 0, sum_pmatdeg = 0;
    bool no_constant_pivot_and_row = true;
    std::vector<int> rest;
    for (int64_t i = 0; i < m; ++i) {
        if (shift[i] < order + diff_shift) {
            pivind[nk++] = i;
        } else {
            if (buf[(size_t)i] == 0 && pdeg[(size_t)(i * m + i)] <= 0)
                no_constant_pivot_and_row = false;
            sum_pmatdeg += buf[(size_t)i];
            rest.push_back((int)i);
        }
    }
    for (size_t t = 0; t < rest.size(); ++t)
        pivind[nk + (int64_t)t] = rest[t];
    // rows[i] <- rows[pivind[i]], shift[i] <- shift[pivind[i]]
    std::vector<int> perm((size_t)m);
    std::vector<int64_t> sh((size_t)m);
    for (int64_t i = 0; i < m; ++i) {
        perm[(size_t)i] = (int)pivind[i];
        sh[(size_t)i] = shift[pivind[i]];
    }
    for (int64_t i = 0; i < m; ++i)
        shift[i] = sh[(size_t)i];
    DMat Kp;
    KTRY(alloc(Kp, m, m, P.len));
    Kp.len = P.len;
    KTRY(gather(X, Kp, P.len, P, 0, m, iota_vec(m), perm));
    P = DMat();

    if (nk >= m - n && no_constant_pivot_and_row) {
        int64_t sum_pivdeg = 0;
        for (int64_t i = 0; i < nk; ++i)
            sum_pivdeg += pdeg[(size_t)(perm[(size_t)i] * m + perm[(size_t)i])];
        if (sum_pivdeg == sum_pmatdeg) {
            K = Kp;
            *nul = nk;
            return PMLG_OK;
        }
    }

    // residual of the other rows, its kernel, and the missing kernel rows
    const DMat approx = rows_of(Kp, nk, m - nk);
    DMat R;
    KTRY(alloc(R, m - nk, n, maxdeg + 1));
    R.len = maxdeg + 1;
    KTRY(mulmid_dev_strided(X.ctx, R.d, R.cap, approx.d, approx.cap, F.d, F.cap, m - nk, m, n,
                            approx.len, F.len, order, order + maxdeg + 1, X.method, X.s));
    std::vector<int64_t> piv2((size_t)(m - nk));
    int64_t nul2 = 0;
    DMat K2;
    KTRY(zls(X, R, shift + nk, piv2.data(), &nul2, K2));
    R = DMat();
    if (nul2 == 0) {
        K = Kp;
        *nul = nk;
        return PMLG_OK;
    }
    DMat prod;
    KTRY(product(X, rows_of(K2, 0, nul2), approx, prod));
    K2 = DMat();
    for (int64_t i = 0; i < nul2; ++i)
        pivind[nk + i] = pivind[nk + piv2[(size_t)i]];
    // merge both runs by increasing pivot index (kernel.c:442-466)
    const int64_t tot = nk + nul2;
    std::vector<int> sel((size_t)tot);
    int64_t i1 = 0, i2 = 0;
    for (int64_t i = 0; i < tot; ++i) {
        if (i2 == nul2 || (i1 < nk && pivind[i1] < pivind[nk + i2]))
            sel[(size_t)i] = (int)i1++;
        else
            sel[(size_t)i] = (int)(nk + i2++);
    }
    DMat Kf;
    const int64_t lf = std::max(Kp.len, prod.len);
    KTRY(alloc(Kf, tot, m, lf));
    Kf.len = lf;
    std::vector<int> d1, s1, d2, s2;
    for (int64_t i = 0; i < tot; ++i) {
        if (sel[(size_t)i] < nk) {
            d1.push_back((int)i);
            s1.push_back(sel[(size_t)i]);
        } else {
            d2.push_back((int)i);
            s2.push_back((int)(sel[(size_t)i] - nk));
        }
    }
    KTRY(gather(X, Kf, lf, Kp, 0, m, d1, s1));
    KTRY(gather(X, Kf, lf, prod, 0, m, d2, s2));
    std::vector<int64_t> pv((size_t)tot), sv((size_t)tot);
    for (int64_t i = 0; i < tot; ++i) {
        pv[(size_t)i] = pivind[sel[(size_t)i]];
        sv[(size_t)i] = shift[sel[(size_t)i]];
    }
    for (int64_t i = 0; i < tot; ++i) {
        pivind[i] = pv[(size_t)i];
        shift[i] = sv[(size_t)i];
    }
    K = Kf;
    *nul = tot;
    return PMLG_OK;
}

// kernel.c:148-228
pmlg_status via_approx(Kx &X, const DMat &Fin, int64_t *shift, int64_t *pivind, int64_t *nul,
                       DMat &K)
{
    const int64_t m = Fin.r, n = Fin.c;
    std::vector<int> deg;
    KTRY(entry_degrees(X, Fin, deg));
    const int64_t d = max_of(deg);
    DMat F = Fin;
    F.len = std::max<int64_t>(1, d + 1);
    if (d == -1) {
        KTRY(identity(X, m, K));
        for (int64_t i = 0; i < m; ++i)
            pivind[i] = i;
        *nul = m;
        return PMLG_OK;
    }
    if (m == 1) {
        *nul = 0;
        return PMLG_OK;
    }
    if (n > kMaxApproxCols)
        return fail(PMLG_ERR_ARG, "kernel basis via approximation: at most 128 columns");
    if (m <= n) {
        // full row rank of the constant coefficient means an empty kernel
        DMat C;
        KTRY(alloc(C, m, n, 1));
        C.len = 1;
        KTRY(gather(X, C, 1, F, 0, n, iota_vec(m), iota_vec(m)));
        DMat scratch;
        KTRY(alloc(scratch, m + 2, 1, 1));
        int64_t rk = 0;
        KTRY(const_rank_dev(X.ctx, C.d, scratch.d, m, n, &rk, X.s));
        if (rk == m) {
            *nul = 0;
            return PMLG_OK;
        }
    }
    int64_t mn = shift[0], mx = shift[0];
    for (int64_t i = 0; i < m; ++i) {
        mn = std::min(mn, shift[i]);
        mx = std::max(mx, shift[i]);
    }
    const int64_t amp = mx - mn;
    const int64_t order = std::min(m, n) * d + amp + d + 1;
    std::vector<int64_t> rdeg(shift, shift + m);
    DMat P;
    KTRY(alloc(P, m, m, order + 1));
    int64_t lenP = 0;
    KTRY(basis_common(X.ctx, P.d, P.cap, &lenP, rdeg.data(), F.d, F.cap, m, n, F.len, order,
                      X.method, false, X.s, nullptr));
    P.len = std::max<int64_t>(lenP, 1);
    KTRY(trim(X, P));
    std::vector<int> rows;
    for (int64_t i = 0; i < m; ++i)
        if (rdeg[(size_t)i] - shift[i] + amp < order - d)
            rows.push_back((int)i);
    const int64_t nk = (int64_t)rows.size();
    for (int64_t t = 0; t < nk; ++t) {
        shift[t] = rdeg[(size_t)rows[(size_t)t]];
        pivind[t] = rows[(size_t)t];
    }
    *nul = nk;
    if (nk == 0)
        return PMLG_OK;
    KTRY(alloc(K, nk, m, P.len));
    K.len = P.len;
    return gather(X, K, P.len, P, 0, m, iota_vec(nk), rows);
}

pmlg_status kernel_run(Kx &X, const uint64_t *dF, int64_t m, int64_t n, int64_t lenF, int algo,
                       int64_t *shift, int64_t *pivind, int64_t *nul, DMat &K)
{
    DMat F;
    if (lenF > 0) {
        F.d = const_cast<uint64_t *>(dF);
        F.r = m;
        F.c = n;
        F.cap = lenF;
        F.len = lenF;
    } else {
        KTRY(alloc(F, m, n, 1));
        F.len = 1;
        HIP_TRY(fill_zero(F.d, m * n, X.s), "fill_zero");
    }
    if (algo == PMLG_KERNEL_VIA_APPROX)
        return via_approx(X, F, shift, pivind, nul, K);
    return zls(X, F, shift, pivind, nul, K);
}

pmlg_status check_kernel_args(pmlg_ctx ctx, int64_t m, int64_t n, int64_t lenF, int algo,
                              const int64_t *shift, const int64_t *pivind, const int64_t *nullity,
                              const int64_t *lenK)
{
    if (m < 1 || m > kMaxRows || n < 1 || n > (1 << 20) || lenF < 0)
        return fail(PMLG_ERR_ARG, "kernel basis: 1 <= rows <= 256, columns >= 1, lenF >= 0");
    if (algo != PMLG_KERNEL_ZLS && algo != PMLG_KERNEL_VIA_APPROX)
        return fail(PMLG_ERR_ARG, "kernel basis: unknown algorithm");
    if (algo == PMLG_KERNEL_VIA_APPROX && n > kMaxApproxCols)
        return fail(PMLG_ERR_ARG, "kernel basis via approximation: at most 128 columns");
    if (!shift || !pivind || !nullity || !lenK)
        return fail(PMLG_ERR_ARG, "kernel basis: null output pointer");
    if (!ctx->is_prime && ctx->p != 2)
        return fail(PMLG_ERR_UNSUPPORTED, "kernel bases need a prime modulus");
    return PMLG_OK;
}

// normalised length of the first nul rows of K (0 when nul == 0)
pmlg_status kernel_length(Kx &X, const DMat &K, int64_t nul, int64_t *len)
{
    if (nul == 0) {
        *len = 0;
        return PMLG_OK;
    }
    std::vector<int> deg;
    KTRY(entry_degrees(X, rows_of(K, 0, nul), deg));
    *len = std::max(1, max_of(deg) + 1);
    return PMLG_OK;
}

}  // namespace

}  // namespace pmlg

using namespace pmlg;

extern "C" {

pmlg_status pmlg_kernel_basis_dev(pmlg_ctx ctx, uint64_t *dK, int64_t capK, int64_t *lenK,
                                  int64_t *nullity, int64_t *pivind, int64_t *shift,
                                  const uint64_t *dF, int64_t m, int64_t n, int64_t lenF, int algo,
                                  int method, void *stream)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    KTRY(check_kernel_args(ctx, m, n, lenF, algo, shift, pivind, nullity, lenK));
    if ((lenF > 0 && !dF) || !dK)
        return fail(PMLG_ERR_ARG, "kernel basis: null buffer");
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    Kx X{ctx, stream ? (hipStream_t)stream : ctx->stream, method, {}};
    int64_t nul = 0;
    DMat K;
    KTRY(kernel_run(X, dF, m, n, lenF, algo, shift, pivind, &nul, K));
    int64_t L = 0;
    KTRY(kernel_length(X, K, nul, &L));
    *nullity = nul;
    *lenK = L;
    if (L > capK)
        return fail(PMLG_ERR_ARG, "kernel basis: output capacity " + std::to_string(capK) +
                                      " below the basis length " + std::to_string(L));
    if (nul > 0) {
        DMat out;
        out.d = dK;
        out.r = m;
        out.c = m;
        out.cap = capK;
        KTRY(gather(X, out, capK, K, 0, m, iota_vec(nul), iota_vec(nul)));
    }
    HIP_TRY(hipStreamSynchronize(X.s), "sync");
    return PMLG_OK;
}

pmlg_status pmlg_kernel_basis(pmlg_ctx ctx, uint64_t *K, int64_t capK, int64_t *lenK,
                              int64_t *nullity, int64_t *pivind, int64_t *shift, const uint64_t *F,
                              int64_t m, int64_t n, int64_t lenF, int algo, int method)
{
    pmlg_status st = check_ctx(ctx);
    if (st != PMLG_OK)
        return st;
    KTRY(check_kernel_args(ctx, m, n, lenF, algo, shift, pivind, nullity, lenK));
    if ((lenF > 0 && !F) || !K)
        return fail(PMLG_ERR_ARG, "kernel basis: null buffer");
    std::lock_guard<std::mutex> lk(ctx->mu);
    DeviceGuard g(ctx->device);
    Kx X{ctx, ctx->stream, method, {}};
    DMat dF;
    if (lenF > 0) {
        // through the pinned staging chunks, entry by entry
        KTRY(alloc(dF, m, n, lenF));
        std::vector<const uint64_t *> ptr((size_t)(m * n));
        std::vector<int64_t> len((size_t)(m * n), lenF);
        for (int64_t e = 0; e < m * n; ++e)
            ptr[(size_t)e] = F + e * lenF;
        KTRY(stage_upload(ctx, dF.d, ptr.data(), len.data(), m * n, lenF));
    }
    int64_t nul = 0;
    DMat Kd;
    KTRY(kernel_run(X, dF.d, m, n, lenF, algo, shift, pivind, &nul, Kd));
    int64_t L = 0;
    KTRY(kernel_length(X, Kd, nul, &L));
    *nullity = nul;
    *lenK = L;
    if (L > capK)
        return fail(PMLG_ERR_ARG, "kernel basis: output capacity " + std::to_string(capK) +
                                      " below the basis length " + std::to_string(L));
    if (nul > 0) {
        // rows [0, nul) packed to the basis length on the device, then read
        // back through the pinned staging chunks, whose unpacking threads
        // zero-fill each entry up to capK
        const int64_t Lw = L > 0 ? L : 1;
        DMat out;
        KTRY(alloc(out, nul, m, Lw));
        KTRY(gather(X, out, Lw, Kd, 0, m, iota_vec(nul), iota_vec(nul)));
        std::vector<uint64_t *> ptr((size_t)(nul * m));
        std::vector<int64_t> len((size_t)(nul * m));
        for (int64_t e = 0; e < nul * m; ++e)
            ptr[(size_t)e] = K + e * capK;
        KTRY(stage_download(ctx, ptr.data(), len.data(), out.d, nul * m, Lw, X.s, capK));
    }
    HIP_TRY(hipStreamSynchronize(X.s), "sync");
    return PMLG_OK;
}

}  // extern "C"
