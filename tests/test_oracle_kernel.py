"""Pins the oracle's kernel bases (oracle/kernel.py) to facts the mathematics
fixes, not to a retyping of its own steps:

- a 2 x 1 column [h a; h b] has the one-row kernel lambda [b, -a] (gcd(a, b)
  = 1 for random a, b): its degrees are exactly (deg b, deg a);
- for a generic m x n matrix of degree d (m > n), a minimal left kernel basis
  has m - n rows whose row degrees add up to n d (the degree of the maximal
  minors of F, which are coprime);
- the two algorithms of the reference (approximant basis at a large order,
  and Zhou-Labahn-Storjohann) must agree on the shifted pivot index and
  pivot degrees, which depend only on the kernel module and the shift;
- the checks of the reference's test (flint-extras/src/nmod_poly_mat_extra/
  test/t-kernel.c:28-100): is_kernel, and the returned shifted row degrees and
  pivot indices equal those recomputed from the basis;
- nullity = m - rank with the rank computed independently by evaluation;
- is_kernel rejects planted mistakes (a dropped row, a perturbed coefficient,
  two rows swapped).
"""
import zlib

import numpy as np
import pytest

from oracle import oracle as O
from oracle import kernel as KO


def _seed(*names):
    # stable across processes (str hashes are salted per process)
    return zlib.crc32("/".join(names).encode())


P60 = 1152921504556515329
P30 = 998244353

SHIFTS = {
    "uniform": lambda m, d: np.zeros(m),
    "increasing": lambda m, d: np.arange(m),
    "decreasing": lambda m, d: m - np.arange(m),
    "shuffle": lambda m, d: np.random.default_rng(m).permutation(m),
    "hermite": lambda m, d: np.arange(m) * m * d,
    "rhermite": lambda m, d: (m - np.arange(m)) * m * d,
    "plateau": lambda m, d: np.where(np.arange(m) < m // 2, 0, m * d),
    "rplateau": lambda m, d: np.where(np.arange(m) < m // 2, m * d, 0),
}


def family(rng, kind, m, n, ln, p):
    if kind == "uniform":
        return O.rand_uniform((m, n, ln), p, int(rng.integers(1 << 30)))
    if kind == "zero":
        return np.zeros((m, n, ln), dtype=np.uint64)
    if kind == "sparse":
        F = O.rand_uniform((m, n, ln), p, int(rng.integers(1 << 30)))
        return F * (rng.random((m, n, ln)) < 0.1)
    if kind == "rdeg":  # unbalanced row degrees
        F = O.rand_uniform((m, n, ln), p, int(rng.integers(1 << 30)))
        for i in range(m):
            F[i, :, int(rng.integers(1, ln + 1)):] = 0
        return F
    if kind == "cdeg":  # unbalanced column degrees
        F = O.rand_uniform((m, n, ln), p, int(rng.integers(1 << 30)))
        for j in range(n):
            F[:, j, int(rng.integers(1, ln + 1)):] = 0
        return F
    if kind == "rkdef":  # rank at most r = max(1, min(m, n) - 1): F = G H
        r = max(1, min(m, n) - 1)
        lg = max(1, (ln + 1) // 2)
        G = O.rand_uniform((m, r, lg), p, int(rng.integers(1 << 30)))
        H = O.rand_uniform((r, n, ln - lg + 1), p, int(rng.integers(1 << 30)))
        return O.polmat_mul(G, H, p)
    raise ValueError(kind)


def check_reference_properties(F, shift, p, nul, K, piv, sh):
    """The checks of t-kernel.c:44-70 plus the nullity."""
    assert KO.is_kernel(K, shift, F, p)
    if nul:
        assert np.array_equal(KO.row_degree(K, shift), sh)
        assert np.array_equal(KO.pivot_index(K, shift)[0], piv)
    assert nul == F.shape[0] - KO.polmat_rank(F, p)


def test_two_by_one_closed_form():
    rng = np.random.default_rng(1)
    for p in (P30, P60):
        for (da, db, dh) in [(3, 5, 2), (7, 1, 0), (0, 4, 3), (6, 6, 1)]:
            a = O.rand_uniform((1, 1, da + 1), p, int(rng.integers(1 << 30)))
            b = O.rand_uniform((1, 1, db + 1), p, int(rng.integers(1 << 30)))
            h = O.rand_uniform((1, 1, dh + 1), p, int(rng.integers(1 << 30)))
            for v in (a, b, h):
                v[0, 0, -1] = v[0, 0, -1] or 1   # exact degrees
            ha, hb = O.polmat_mul(h, a, p), O.polmat_mul(h, b, p)
            ln = max(ha.shape[2], hb.shape[2])
            F = np.zeros((2, 1, ln), dtype=np.uint64)
            F[0, :, :ha.shape[2]] = ha[0]
            F[1, :, :hb.shape[2]] = hb[0]
            for shift in ([0, 0], [0, 50], [9, 0]):
                for fn in (KO.kernel_via_approx, KO.kernel_zls):
                    nul, K, piv, sh = fn(F, shift, p)
                    assert nul == 1
                    assert not O.polmat_mul(K, F, p).any()
                    assert list(KO.degrees(K)[0]) == [db, da]


@pytest.mark.parametrize("m,n,d", [(4, 1, 6), (5, 2, 4), (7, 3, 3), (9, 4, 2), (6, 5, 3)])
def test_generic_degree_sum(m, n, d):
    rng = np.random.default_rng(m * 100 + n)
    F = O.rand_uniform((m, n, d + 1), P60, int(rng.integers(1 << 30)))
    zero = np.zeros(m, dtype=np.int64)
    for fn in (KO.kernel_via_approx, KO.kernel_zls):
        nul, K, piv, sh = fn(F, zero, P60)
        assert nul == m - n
        assert int(sh.sum()) == n * d
        assert not O.polmat_mul(K, F, P60).any()
        # irreducible: full row rank at a random point
        assert O.rank(O.polmat_eval(K, 123456789, P60), P60) == nul


@pytest.mark.parametrize("shift_name", sorted(SHIFTS))
@pytest.mark.parametrize("kind", ["uniform", "zero", "sparse", "rdeg", "cdeg", "rkdef"])
def test_two_algorithms_agree(shift_name, kind):
    rng = np.random.default_rng(_seed(shift_name, kind))
    for it in range(3):
        p = (P30, P60)[it % 2]
        m, n = int(rng.integers(2, 8)), int(rng.integers(1, 6))
        ln = int(rng.integers(1, 7))
        F = family(rng, kind, m, n, ln, p)
        shift = SHIFTS[shift_name](m, ln).astype(np.int64)
        r1 = KO.kernel_via_approx(F, shift, p)
        r2 = KO.kernel_zls(F, shift, p)
        for (nul, K, piv, sh) in (r1, r2):
            check_reference_properties(F, shift, p, nul, K, piv, sh)
        assert r1[0] == r2[0]
        assert np.array_equal(r1[2], r2[2])
        assert np.array_equal(KO.pivot_index(r1[1], shift)[1], KO.pivot_index(r2[1], shift)[1])


@pytest.mark.parametrize("orient", [KO.ROW_LOWER, KO.ROW_UPPER, KO.COL_LOWER, KO.COL_UPPER])
def test_orientations(orient):
    rng = np.random.default_rng(5)
    for it in range(4):
        m, n = int(rng.integers(1, 7)), int(rng.integers(1, 7))
        F = family(rng, ("uniform", "rkdef")[it % 2], m, n, int(rng.integers(1, 5)), P30)
        left = orient in (KO.ROW_LOWER, KO.ROW_UPPER)
        dim = m if left else n
        shift = rng.integers(0, 4, dim).astype(np.int64)
        nul, K, piv, sh = KO.kernel(F, shift, P30, orient)
        assert KO.is_kernel(K, shift, F, P30, orient)
        if nul:
            degs = KO.row_degree(K, shift) if left else KO.column_degree(K, shift)
            assert np.array_equal(degs, sh)
            assert np.array_equal(KO.pivot_index(K, shift, orient)[0], piv)
        assert nul == dim - KO.polmat_rank(F, P30)


def test_is_kernel_rejects_planted_mistakes():
    rng = np.random.default_rng(9)
    F = family(rng, "uniform", 6, 2, 4, P60)
    shift = np.zeros(6, dtype=np.int64)
    nul, K, piv, sh = KO.kernel_zls(F, shift, P60)
    assert nul == 4 and KO.is_kernel(K, shift, F, P60)
    assert not KO.is_kernel(K[:-1], shift, F, P60)            # a row missing
    bad = K.copy()
    bad[1, 2, 0] = (int(bad[1, 2, 0]) + 1) % P60
    assert not KO.is_kernel(bad, shift, F, P60)               # not in the kernel
    assert not KO.is_kernel(K[[1, 0, 2, 3]], shift, F, P60)   # pivots not increasing


def test_polmat_rank_against_construction():
    rng = np.random.default_rng(11)
    for r in (0, 1, 2, 3):
        G = O.rand_uniform((5, max(r, 1), 3), P30, 7 + r)
        H = O.rand_uniform((max(r, 1), 4, 2), P30, 17 + r)
        F = O.polmat_mul(G, H, P30) if r else np.zeros((5, 4, 4), dtype=np.uint64)
        assert KO.polmat_rank(F, P30) == r
