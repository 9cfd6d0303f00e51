"""Closed forms that pin the oracle's exact approximant and kernel bases.

The reference cannot run here, so the oracle's M-Basis / PM-Basis output is
compared bit for bit with bases computed from the mathematics of an M-Basis
step alone, written here in a few lines of Python integer arithmetic (no
oracle code, no numpy arithmetic), following the step as
flint-extras/src/nmod_mat_poly_extra/nmod_mat_poly_mbasis.c:116-378 defines it:

* the residual of the current basis P at order d is coefficient d of P F;
* the rows are ranked by (shift, index), stably;
* the left kernel of the ranked residual is read from the reduced row echelon
  form of its transpose (left_nullspace.c:42-102): the pivot columns of that
  form are the residual's row rank profile, and the kernel vector of a
  non-pivot row j is e_j minus the echelon form's column j spread over the
  pivot rows;
* each non-pivot row becomes its kernel combination of the current rows, the
  pivot rows are multiplied by x and their shifts grow by one.

Three cases are checked against the oracle:

(i)   order 1 (one step) with non-uniform shifts that include ties, from the
      echelon form computed here;
(ii)  one column (n = 1), orders 1..40, random shifts: the step reduces to
      "the pivot is the nonzero residual row of least (shift, index); every
      other row i becomes row_i - (r_i / r_piv) row_piv; the pivot row is
      multiplied by x and its shift gains 1", against O.mbasis and O.pmbasis;
(iii) the one-row kernel of a 2 x 1 matrix through kernel_via_approx
      (kernel.c:148-228: one approximant basis at order
      min(m, n) d + amp + d + 1, then the rows with rdeg_s - s + amp below
      order - d), from (ii); the ZLS kernel is the same row up to a nonzero
      constant (a rank-one kernel basis is unique up to a unit).

A general-width version of the same steps (several columns, several orders)
is also compared with the oracle, covering the composed rankings and updates
over many steps.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import kernel as KO

P30 = 998244353
P60 = 1152921504556515329


def _trim(a):
    a = list(a)
    while a and a[-1] == 0:
        a.pop()
    return a


def _coeff_of_product(row, F, j, d):
    """coefficient d of sum_t row[t] * F[t][j] (polynomials as int lists)"""
    c = 0
    for t, a in enumerate(row):
        b = F[t][j]
        lo, hi = max(0, d - len(b) + 1), min(d, len(a) - 1)
        for u in range(lo, hi + 1):
            c += a[u] * b[d - u]
    return c


def _axpy(y, q, x, p):
    """y + q x for polynomials (q a residue)"""
    out = list(y) + [0] * max(0, len(x) - len(y))
    for u, v in enumerate(x):
        out[u] = (out[u] + q * v) % p
    return _trim(out)


def _rref_transpose(R, p):
    """Reduced row echelon form of R^T (R: m rows of n residues), by plain
    Gauss-Jordan with the leftmost available pivot; returns (rows, pivot columns)."""
    m, n = len(R), len(R[0]) if R else 0
    T = [[R[i][j] % p for i in range(m)] for j in range(n)]
    piv, r = [], 0
    for col in range(m):
        pr = next((x for x in range(r, n) if T[x][col]), None)
        if pr is None:
            continue
        T[r], T[pr] = T[pr], T[r]
        inv = pow(T[r][col], p - 2, p)
        T[r] = [v * inv % p for v in T[r]]
        for x in range(n):
            if x != r and T[x][col]:
                f = T[x][col]
                T[x] = [(a - f * b) % p for a, b in zip(T[x], T[r])]
        piv.append(col)
        r += 1
    return T[:r], piv


def closed_mbasis(F, order, shift, p):
    """M-Basis from its step description only.  F: m x n list of int
    polynomials.  Returns (P as m x m polynomial lists, shifts)."""
    m, n = len(F), len(F[0])
    P = [[[1] if i == j else [] for j in range(m)] for i in range(m)]
    s = [int(x) for x in shift]
    for d in range(order):
        res = [[_coeff_of_product(P[i], F, j, d) % p for j in range(n)] for i in range(m)]
        rank_order = sorted(range(m), key=lambda i: (s[i], i))
        E, pivcols = _rref_transpose([res[i] for i in rank_order], p)
        if len(pivcols) == 0:
            continue  # zero residual: every row is already an approximant at order d + 1
        nonpiv = [c for c in range(m) if c not in pivcols]
        newP = [list(map(list, row)) for row in P]
        for c in nonpiv:
            i = rank_order[c]
            for r, pc in enumerate(pivcols):
                q = (-E[r][c]) % p
                if q:
                    k = rank_order[pc]
                    newP[i] = [_axpy(newP[i][j], q, P[k][j], p) for j in range(m)]
        for pc in pivcols:
            k = rank_order[pc]
            newP[k] = [([0] + e) if e else [] for e in P[k]]
            s[k] += 1
        P = newP
    return P, s


def closed_mbasis_one_column(f, order, shift, p):
    """(ii): the n = 1 rule as stated, independent of the general version."""
    m = len(f)
    P = [[[1] if i == j else [] for j in range(m)] for i in range(m)]
    s = [int(x) for x in shift]
    for d in range(order):
        r = [_coeff_of_product(P[i], [[g] for g in f], 0, d) % p for i in range(m)]
        nz = [i for i in range(m) if r[i]]
        if not nz:
            continue
        c = min(nz, key=lambda i: (s[i], i))
        inv = pow(r[c], p - 2, p)
        for i in range(m):
            if i != c and r[i]:
                q = (-r[i] * inv) % p
                P[i] = [_axpy(P[i][j], q, P[c][j], p) for j in range(m)]
        P[c] = [([0] + e) if e else [] for e in P[c]]
        s[c] += 1
    return P, s


def to_lists(F):
    return [[_trim(int(v) for v in F[i, j]) for j in range(F.shape[1])] for i in range(F.shape[0])]


def to_array(P):
    m = len(P)
    L = max([len(e) for row in P for e in row] + [1])
    A = np.zeros((m, m, L), dtype=np.uint64)
    for i in range(m):
        for j in range(m):
            A[i, j, :len(P[i][j])] = P[i][j]
    return A


def same(A, B):
    return np.array_equal(O.normalise(A), O.normalise(B))


def test_order_one_shifted_with_ties():
    """(i): one step, shifts with ties and distinct values, several ranks."""
    rng = np.random.default_rng(31)
    for p in (P60, P30, 7):
        for (m, n) in [(5, 2), (6, 3), (4, 4), (7, 1), (3, 5)]:
            for trial in range(4):
                F = rng.integers(0, p, size=(m, n, 2), dtype=np.uint64)
                if trial == 1:
                    F[1] = F[0]          # repeated rows: rank below min(m, n)
                if trial == 2:
                    F[:, 0] = 0          # a zero column
                shift = rng.integers(0, 3, size=m)  # ties among 0, 1, 2
                if trial == 3:
                    shift = np.array([5, -1, 5, 2, -1, 0, 7][:m])
                # the step, written out: rank rows by (shift, index), echelon
                # form of the ranked residual's transpose, pivots and kernel rows
                res = [[int(v) for v in F[i, :, 0]] for i in range(m)]
                ranked = sorted(range(m), key=lambda i: (int(shift[i]), i))
                E, pivcols = _rref_transpose([res[i] for i in ranked], p)
                exp = np.zeros((m, m, 2), dtype=np.uint64)
                exp_s = [int(x) for x in shift]
                nonpiv = [c for c in range(m) if c not in pivcols]
                for pc in pivcols:
                    k = ranked[pc]
                    exp[k, k, 1] = 1
                    exp_s[k] += 1
                for c in nonpiv:
                    i = ranked[c]
                    exp[i, i, 0] = 1
                    for r, pc in enumerate(pivcols):
                        exp[i, ranked[pc], 0] = (-E[r][c]) % p
                for fn in (O.mbasis, O.pmbasis):
                    P, sh = fn(F, 1, shift, p)
                    assert same(P, exp), (p, m, n, trial, fn.__name__)
                    assert list(sh) == exp_s


def test_one_column_rule_matches_oracle():
    """(ii): n = 1, orders 1..40, random shifts (with ties), both oracle
    algorithms, bit for bit, and the general-width steps agree too."""
    rng = np.random.default_rng(32)
    for it in range(40):
        p = (P30, P60, 101, 5)[it % 4]
        m = int(rng.integers(1, 7))
        order = 1 + it
        ln = int(rng.integers(1, order + 3))
        F = rng.integers(0, p, size=(m, 1, ln), dtype=np.uint64)
        if it % 5 == 0:
            F[0] = 0                       # a zero row
        if it % 7 == 0 and m > 1:
            F[m - 1] = F[0]                # a repeated row
        shift = rng.integers(-3, 4, size=m)
        P, s = closed_mbasis_one_column([[int(v) for v in F[i, 0]] for i in range(m)], order,
                                        shift, p)
        exp = to_array(P)
        for fn in (O.mbasis, O.pmbasis):
            Po, so = fn(F, order, shift, p)
            assert same(Po, exp), (it, fn.__name__)
            assert list(so) == s, (it, fn.__name__)
        Pg, sg = closed_mbasis(to_lists(F), order, shift, p)
        assert same(to_array(Pg), exp) and sg == s


@pytest.mark.parametrize("m,n,order", [(3, 2, 12), (4, 2, 20), (5, 3, 16), (6, 4, 10), (4, 4, 9)])
def test_general_steps_match_oracle(m, n, order):
    """Several columns, several orders: the composed rankings, the kernel
    combinations and the x-shifts over every step, including low primes
    (more rank drops) and non-uniform shifts."""
    rng = np.random.default_rng(m * 10 + n)
    for p in (P60, P30, 3):
        for trial in range(3):
            F = rng.integers(0, p, size=(m, n, order), dtype=np.uint64)
            shift = rng.integers(0, 4, size=m) if trial else np.zeros(m, dtype=np.int64)
            P, s = closed_mbasis(to_lists(F), order, shift, p)
            for fn in (O.mbasis, O.pmbasis):
                Po, so = fn(F, order, shift, p)
                assert same(Po, to_array(P)), (p, trial, fn.__name__)
                assert list(so) == s


def test_two_by_one_kernel_row_exact():
    """(iii): the exact kernel row of a 2 x 1 matrix from kernel_via_approx's
    definition applied to the closed-form basis of (ii)."""
    rng = np.random.default_rng(33)
    for p in (P30, P60):
        for (da, db, dh) in [(3, 5, 2), (7, 1, 0), (0, 4, 3), (6, 6, 1), (2, 2, 0)]:
            a = [int(x) for x in rng.integers(0, p, size=da + 1)]
            b = [int(x) for x in rng.integers(0, p, size=db + 1)]
            h = [int(x) for x in rng.integers(0, p, size=dh + 1)]
            a[-1] = a[-1] or 1
            b[-1] = b[-1] or 1
            h[-1] = h[-1] or 1

            def mul(x, y):
                out = [0] * (len(x) + len(y) - 1)
                for i, u in enumerate(x):
                    for j, v in enumerate(y):
                        out[i + j] = (out[i + j] + u * v) % p
                return out
            f = [mul(h, a), mul(h, b)]
            ln = max(len(f[0]), len(f[1]))
            F = np.zeros((2, 1, ln), dtype=np.uint64)
            F[0, 0, :len(f[0])] = f[0]
            F[1, 0, :len(f[1])] = f[1]
            d = ln - 1
            for shift in ([0, 0], [0, 50], [9, 0], [3, 3]):
                amp = max(shift) - min(shift)
                order = min(2, 1) * d + amp + d + 1
                P, s = closed_mbasis_one_column(f, order, shift, p)
                rows = [i for i in range(2) if s[i] - shift[i] + amp < order - d]
                assert len(rows) == 1
                row = P[rows[0]]
                L = max(len(e) for e in row)
                exp = np.zeros((1, 2, L), dtype=np.uint64)
                for j in range(2):
                    exp[0, j, :len(row[j])] = row[j]
                nul, K, piv, sh = KO.kernel_via_approx(F, shift, p)
                assert nul == 1 and list(piv) == rows and list(sh) == [s[rows[0]]]
                assert same(K, exp), (p, da, db, dh, shift)
                # ZLS: the same row up to a nonzero constant
                nul2, K2, piv2, sh2 = KO.kernel_zls(F, shift, p)
                assert nul2 == 1 and list(piv2) == list(piv) and list(sh2) == list(sh)
                K2n, Kn = O.normalise(K2), O.normalise(K)
                assert K2n.shape == Kn.shape
                j0 = int(np.nonzero(Kn[0, :, :].reshape(-1))[0][0])
                lam = int(K2n.reshape(-1)[j0]) * pow(int(Kn.reshape(-1)[j0]), p - 2, p) % p
                assert lam != 0
                assert np.array_equal(K2n.astype(object), (Kn.astype(object) * lam) % p)
