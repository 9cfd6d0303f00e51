"""CPU restatement of PML's minimal kernel bases over Z/pZ.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
package never imports this module.

Follows flint-extras/src/nmod_poly_mat_extra/kernel.c:
  kernel_via_approx  <- nmod_poly_mat_kernel_via_approx   (kernel.c:148-228)
  kernel_zls         <- nmod_poly_mat_kernel_zls_approx   (kernel.c:232-478)
  kernel             <- nmod_poly_mat_kernel, 4 orientations (kernel.c:26-146)
and the checks it is tested with:
  pivot_index        <- _nmod_poly_vec_pivot_profile      (degrees_pivots_leadingmatrix.c:276-365)
  row_degree         <- nmod_poly_mat_row_degree          (degrees_pivots_leadingmatrix.c:23-75)
  is_ordered_weak_popov <- nmod_poly_mat_is_form.c:89-123
  is_kernel          <- nmod_poly_mat_is_kernel           (verification.c:119-215)
  polmat_rank        <- FLINT's nmod_poly_mat_rank, computed here exactly by
                        evaluation: a nonzero r x r minor has degree at most
                        r*deg(F), so it is nonzero at one of r*deg(F)+1 points.

The approximant bases, products and middle products come from the C oracle
(oracle.py), so this file is host bookkeeping only.  Matrices are numpy uint64
arrays (rows, cols, len); shifts and pivot indices are int64 arrays.
"""
import numpy as np

from oracle import oracle as O

ROW_LOWER, ROW_UPPER, COL_LOWER, COL_UPPER = "row_lower", "row_upper", "col_lower", "col_upper"


def degrees(A):
    """Entry degrees (-1 for the zero polynomial)."""
    A = np.asarray(A)
    r, c, ln = A.shape
    if ln == 0:
        return np.full((r, c), -1, dtype=np.int64)
    nz = A != 0
    last = ln - 1 - np.argmax(nz[..., ::-1], axis=2)
    return np.where(nz.any(axis=2), last, -1).astype(np.int64)


def identity(m):
    K = np.zeros((m, m, 1), dtype=np.uint64)
    for i in range(m):
        K[i, i, 0] = 1
    return K


def _cdiv(a, b):
    """C integer division (truncation toward zero), as in kernel.c:352."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b > 0) else -q


def _pad(A, ln):
    if A.shape[2] >= ln:
        return A
    out = np.zeros(A.shape[:2] + (ln,), dtype=np.uint64)
    out[:, :, :A.shape[2]] = A
    return out


def kernel_via_approx(F, shift, p):
    """kernel.c:148-228.  Returns (nullity, K, pivind, shift_out): K holds
    the nullity kernel rows, shift_out their shifted row degrees."""
    F = np.ascontiguousarray(F, dtype=np.uint64)
    m, n, _ = F.shape
    shift = np.array(shift, dtype=np.int64)
    deg = degrees(F)
    d = int(deg.max()) if deg.size else -1
    if d == -1:
        return m, identity(m), np.arange(m, dtype=np.int64), shift.copy()
    if m == 1:
        return 0, np.zeros((0, m, 1), dtype=np.uint64), np.zeros(0, np.int64), np.zeros(0, np.int64)
    if m <= n and O.rank(np.ascontiguousarray(F[:, :, 0]), p) == m:
        return 0, np.zeros((0, m, 1), dtype=np.uint64), np.zeros(0, np.int64), np.zeros(0, np.int64)
    amp = int(shift.max() - shift.min())
    order = min(m, n) * d + amp + d + 1
    P, rdeg = O.pmbasis(F, order, shift, p)
    rows = [i for i in range(m) if rdeg[i] - shift[i] + amp < order - d]
    return (len(rows), P[rows], np.array(rows, dtype=np.int64), rdeg[rows].copy())


def _zls(F, shift, pivind, p):
    """kernel.c:232-478 on views: shift (length m, in/out) and pivind (length
    m, out) are modified in place.  Returns (nullity, K) with K having at
    least nullity rows, the first nullity forming the basis."""
    m, n, lf = F.shape
    deg = degrees(F)
    if m == 0 or n == 0 or ((m == 1 or n == 1) and not (deg >= 0).any()):
        pivind[:m] = np.arange(m)
        return m, identity(m)
    if m == 1:
        return 0, np.zeros((1, 1, 1), dtype=np.uint64)

    if 2 * n > m:
        # split F = [F1 | F2], F1 with n - n//2 columns (kernel.c:252-313)
        n1 = n - n // 2
        nul, K = _zls(np.ascontiguousarray(F[:, :n1]), shift, pivind, p)
        if nul == 0:
            return 0, K
        ker1 = np.ascontiguousarray(K[:nul])
        R = O.polmat_mul(ker1, np.ascontiguousarray(F[:, n1:]), p)
        piv2 = np.zeros(nul, dtype=np.int64)
        nul2, K2 = _zls(R, shift[:nul], piv2, p)
        if nul2 == 0:
            return 0, K2
        K3 = O.polmat_mul(np.ascontiguousarray(K2[:nul2]), ker1, p)
        for i in range(nul2):
            pivind[i] = pivind[piv2[i]]
        return nul2, K3

    # 1 <= n <= m/2 (kernel.c:315-477)
    rmax = deg.max(axis=1)
    buf = np.where(rmax >= 0, rmax, 0)
    if not (rmax >= 0).any():
        pivind[:m] = np.arange(m)
        return m, identity(m)
    maxdeg = int(buf.max())
    diff_shift = int(min(int(shift[i]) - int(buf[i]) for i in range(m)))
    num = -m * diff_shift - 1 + int(sum(int(s) for s in shift[:m]))
    order = 1 + _cdiv(num, m - n)
    order = 1 + max(maxdeg, order)

    P, s_out = O.pmbasis(F, order, shift[:m], p)
    shift[:m] = s_out
    pdeg = degrees(P)

    nul = 0
    rest = []
    sum_pmatdeg = 0
    no_constant_pivot_and_row = True
    for i in range(m):
        if shift[i] < order + diff_shift:
            pivind[nul] = i
            nul += 1
        else:
            if buf[i] == 0 and pdeg[i, i] <= 0:
                no_constant_pivot_and_row = False
            sum_pmatdeg += int(buf[i])
            rest.append(i)
    pivind[nul:m] = rest
    perm = pivind[:m].copy()
    K = P[perm]
    shift[:m] = shift[:m][perm]

    if nul >= m - n and no_constant_pivot_and_row:
        sum_pivdeg = sum(int(pdeg[perm[i], perm[i]]) for i in range(nul))
        if sum_pivdeg == sum_pmatdeg:
            return nul, K

    approx = np.ascontiguousarray(K[nul:])
    R = O.polmat_mulmid(approx, F, order, order + maxdeg + 1, p)
    piv2 = np.zeros(m - nul, dtype=np.int64)
    nul2, K2 = _zls(R, shift[nul:m], piv2, p)
    if nul2 == 0:
        return nul, K

    prod = O.polmat_mul(np.ascontiguousarray(K2[:nul2]), approx, p)
    for i in range(nul2):
        pivind[nul + i] = pivind[nul + piv2[i]]
    # merge the two runs of rows by increasing pivot index (kernel.c:442-466)
    order_rows = []
    i1 = i2 = 0
    for _ in range(nul + nul2):
        if i2 == nul2 or (i1 < nul and pivind[i1] < pivind[nul + i2]):
            order_rows.append(i1)
            i1 += 1
        else:
            order_rows.append(nul + i2)
            i2 += 1
    ln = max(K.shape[2], prod.shape[2])
    rows = np.concatenate([_pad(K[:nul], ln), _pad(prod, ln)], axis=0)
    sel = np.array(order_rows, dtype=np.int64)
    pivind[:nul + nul2] = pivind[:nul + nul2][sel]
    shift[:nul + nul2] = shift[:nul + nul2][sel]
    return nul + nul2, rows[sel]


def kernel_zls(F, shift, p):
    """kernel.c:232-478.  Returns (nullity, K, pivind, shift_out)."""
    F = np.ascontiguousarray(F, dtype=np.uint64)
    m = F.shape[0]
    sh = np.array(shift, dtype=np.int64)
    piv = np.zeros(m, dtype=np.int64)
    nul, K = _zls(F, sh, piv, p)
    if nul == 0:
        K = np.zeros((0, m, 1), dtype=np.uint64)
    return nul, np.ascontiguousarray(K[:nul]), piv[:nul].copy(), sh[:nul].copy()


def _invert_rows(A):
    return np.ascontiguousarray(A[::-1])


def _invert_cols(A):
    return np.ascontiguousarray(A[:, ::-1])


def _transpose(A):
    return np.ascontiguousarray(A.transpose(1, 0, 2))


def kernel(F, shift, p, orient=ROW_LOWER):
    """kernel.c:26-146.  Left kernels (ROW_*) return K with nullity rows,
    right kernels (COL_*) return K with nullity columns.  shift=None uses the
    row (left) or column (right) degrees of F with zero lines taken as 0."""
    F = np.ascontiguousarray(F, dtype=np.uint64)
    deg = degrees(F)
    left = orient in (ROW_LOWER, ROW_UPPER)
    if shift is None:
        d = deg.max(axis=1) if left else deg.max(axis=0)
        shift = np.where(d >= 0, d, 0)
    sh = np.array(shift, dtype=np.int64)
    if orient == ROW_LOWER:
        return kernel_zls(F, sh, p)
    if orient == COL_UPPER:
        nul, K, piv, s = kernel_zls(_transpose(F), sh, p)
        return nul, _transpose(K), piv, s
    # ROW_UPPER / COL_LOWER: mirror the rows (and the shift), solve ROW_LOWER,
    # mirror rows and columns of the result and the pivot indices back
    G = F if orient == ROW_UPPER else _transpose(F)
    dim = G.shape[0]
    nul, K, piv, s = kernel_zls(_invert_rows(G), sh[::-1], p)
    K = _invert_cols(_invert_rows(K))
    s = s[::-1].copy()
    piv = (dim - 1 - piv[::-1]).astype(np.int64)
    if orient == COL_LOWER:
        K = _transpose(K)
    return nul, K, piv, s


def pivot_index(K, shift, orient=ROW_LOWER):
    """Shifted pivot index and pivot degree of every row (ROW_*) or column
    (COL_*), as _nmod_poly_vec_pivot_profile with a shift: ROW_LOWER/COL_UPPER
    take the last entry reaching the shifted degree, ROW_UPPER/COL_LOWER the
    first; a zero vector has index -1 and degree -1 (ROW_LOWER/COL_UPPER) or
    the vector length (ROW_UPPER/COL_LOWER)."""
    K = np.asarray(K)
    deg = degrees(K)
    if orient in (COL_LOWER, COL_UPPER):
        deg = deg.T
    last = orient in (ROW_LOWER, COL_UPPER)
    piv, pdeg = [], []
    for row in deg:
        best, bi = None, -1
        for j, d in enumerate(row):
            if d < 0:
                continue
            v = int(d) + int(shift[j])
            if bi == -1 or (v >= best if last else v > best):
                best, bi = v, j
        piv.append(bi)
        pdeg.append(-1 if bi == -1 and last else (len(row) if bi == -1 else best - int(shift[bi])))
    return np.array(piv, dtype=np.int64), np.array(pdeg, dtype=np.int64)


def row_degree(K, shift):
    """Shifted row degrees; a zero row gets min(shift) - 1."""
    deg = degrees(K)
    mn = int(min(shift)) if len(shift) else 0
    out = []
    for row in deg:
        vals = [int(d) + int(shift[j]) for j, d in enumerate(row) if d >= 0]
        out.append(max(vals) if vals else mn - 1)
    return np.array(out, dtype=np.int64)


def column_degree(K, shift):
    return row_degree(_transpose(np.asarray(K)), shift)


def is_ordered_weak_popov(K, shift, orient=ROW_LOWER):
    piv, _ = pivot_index(K, shift, orient)
    dim = len(piv)
    zpiv = K.shape[1] if orient in (ROW_LOWER, ROW_UPPER) else K.shape[0]
    if dim and piv[0] == -1:
        return False
    if any(piv[i] >= piv[i + 1] for i in range(dim - 1)):
        return False
    if dim and piv[-1] == zpiv:
        return False
    return True


def polmat_rank(F, p):
    """Exact rank over Z/p(x) by evaluation at 0, 1, ..., min(m,n)*deg(F)."""
    F = np.ascontiguousarray(F, dtype=np.uint64)
    m, n, _ = F.shape
    d = int(degrees(F).max()) if m * n else -1
    if d < 0:
        return 0
    full = min(m, n)
    assert full * d < p, "not enough evaluation points in Z/pZ"
    best = 0
    for x in range(full * d + 1):
        best = max(best, O.rank(O.polmat_eval(F, x, p), p))
        if best == full:
            break
    return best


def is_kernel(K, shift, F, p, orient=ROW_LOWER):
    """verification.c:119-215: dimensions, rank, ordered weak Popov form and
    K*F == 0 (left) or F*K == 0 (right).  Generation is not checked (as in
    the reference)."""
    K = np.ascontiguousarray(K, dtype=np.uint64)
    F = np.ascontiguousarray(F, dtype=np.uint64)
    m, n, _ = F.shape
    rk = polmat_rank(F, p)
    if orient in (ROW_LOWER, ROW_UPPER):
        if K.shape[1] != m or K.shape[0] != m - rk:
            return False
        if not is_ordered_weak_popov(K, shift, orient):
            return False
        if K.shape[0] == 0:
            return True
        R = O.polmat_mul(K, F, p) if K.shape[2] and F.shape[2] else np.zeros((1,), np.uint64)
    else:
        if K.shape[0] != n or K.shape[1] != n - rk:
            return False
        if not is_ordered_weak_popov(K, shift, orient):
            return False
        if K.shape[1] == 0:
            return True
        R = O.polmat_mul(F, K, p) if K.shape[2] and F.shape[2] else np.zeros((1,), np.uint64)
    return not R.any()
