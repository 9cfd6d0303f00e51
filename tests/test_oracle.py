"""Pins the CPU oracle (oracle/pml_oracle.c) without the reference.

The reference (flint-extras) needs FLINT >= 3.4 and cannot be built here, and
it ships no golden vectors for this path, so each oracle function is checked
against what the mathematics fixes, chosen so that a plausible mistake (a
dropped term, a wrong index or sign, a transposed operand) fails:

* scalar arithmetic: Python's exact big integers and sympy's primality test;
* product: a pure-Python double loop on small cases, schoolbook vs NTT, and
  the evaluation identity C(r) = A(r) B(r) at random r;
* middle product: the definition used by the reference test
  (test/t-mulmid.c:26-34: multiply, shift right by nlo, truncate to nhi-nlo);
* RREF / left nullspace: the defining properties of a reduced row echelon
  form (which make it unique) and of the compact nullspace convention of
  left_nullspace.c:42-102 and nmod_mat_extra.h:510-527;
* M-Basis / PM-Basis: the certificate of verification.c:21-116 (itself pinned
  by planted mistakes below), the closed form at order 1 (a constant left
  kernel), zero input -> identity, and M-Basis == PM-Basis on generic inputs.
"""
import itertools

import numpy as np
import pytest
import sympy

from oracle import oracle as O

PRIMES = [2, 3, 5, 521, 524309, 536870923, 998244353, 549755813911, 562949953421381,
          576460752303423619, 1152921504556515329, 18446744073709551557]
FFT_PRIMES = [998244353, 7340033, 1152921504556515329, 469762049]


def rand(rng, shape, p):
    return rng.integers(0, p, size=shape, dtype=np.uint64)


# ---------------------------------------------------------------- scalars
def test_scalar_arithmetic():
    rng = np.random.default_rng(0)
    for p in PRIMES:
        for _ in range(50):
            a, b = int(rng.integers(0, p, dtype=np.uint64)), int(rng.integers(0, p, dtype=np.uint64))
            e = int(rng.integers(0, 2**63, dtype=np.uint64))
            assert O.mulmod(a, b, p) == a * b % p
            assert O.powmod(a, e, p) == pow(a, e, p)


def test_primality_matches_sympy():
    rng = np.random.default_rng(1)
    cands = [int(x) for x in rng.integers(2, 2**63, size=300, dtype=np.uint64)]
    cands += [int(x) | 1 for x in rng.integers(2, 2**40, size=300, dtype=np.uint64)]
    cands += PRIMES + [3215031751, 2152302898747, 3474749660383, 341550071728321,
                       3825123056546413051, 318665857834031151167461]
    for n in cands:
        if n < 2**64:
            assert O.is_prime(n) == sympy.isprime(n), n


@pytest.mark.parametrize("p", FFT_PRIMES)
def test_root_of_unity(p):
    v = ((p - 1) & -(p - 1)).bit_length() - 1
    for k in range(0, v + 1):
        r = O.root_of_unity_2k(p, k)
        assert pow(r, 2**k, p) == 1
        if k:
            assert pow(r, 2**(k - 1), p) == p - 1
    assert O.root_of_unity_2k(p, v + 1) == 0


# ---------------------------------------------------------------- product
def brute_mul(A, B, p):
    m, k, la = A.shape
    _, n, lb = B.shape
    C = [[[0] * (la + lb - 1) for _ in range(n)] for _ in range(m)]
    for i, j, kk in itertools.product(range(m), range(n), range(k)):
        for a in range(la):
            x = int(A[i, kk, a])
            if x:
                for b in range(lb):
                    C[i][j][a + b] += x * int(B[kk, j, b])
    return np.array([[[c % p for c in row] for row in r] for r in C], dtype=np.uint64)


@pytest.mark.parametrize("p", PRIMES)
def test_school_matches_brute_force(p):
    rng = np.random.default_rng(p % 1000)
    for (m, k, n, la, lb) in [(1, 1, 1, 1, 1), (2, 3, 4, 5, 2), (3, 1, 2, 7, 9), (2, 5, 1, 1, 6)]:
        A, B = rand(rng, (m, k, la), p), rand(rng, (k, n, lb), p)
        assert np.array_equal(O.polmat_mul(A, B, p, "school"), brute_mul(A, B, p))


def test_school_extreme_values():
    for p in (18446744073709551557, 1152921504556515329, 2):
        A = np.full((2, 40, 30), p - 1, dtype=np.uint64)
        B = np.full((40, 3, 20), p - 1, dtype=np.uint64)
        C = O.polmat_mul(A, B, p, "school")
        # (p-1)^2 = 1 mod p; coefficient d counts the pairs (kk, a, b) with a+b = d
        for d in (0, 5, 19, 29, 48):
            cnt = 40 * sum(1 for a in range(30) if 0 <= d - a < 20)
            assert int(C[0, 0, d]) == cnt % p


@pytest.mark.parametrize("p", FFT_PRIMES + [4611686018429485057])
def test_ntt_matches_school(p):
    """The NTT product (lazy butterflies below 2^62, reduced ones above;
    pointwise panels of 64, 32, ... points; a fold of the 128-bit sums every
    few terms for primes above 2^62) equals the closed form, with one thread
    and with several."""
    rng = np.random.default_rng(p % 997)
    for (m, k, n, la, lb) in [(1, 1, 1, 1, 1), (3, 4, 2, 33, 31), (2, 2, 2, 200, 57),
                              (1, 3, 1, 512, 513), (2, 9, 3, 40, 3), (40, 40, 41, 4, 3)]:
        A, B = rand(rng, (m, k, la), p), rand(rng, (k, n, lb), p)
        ref = O.polmat_mul(A, B, p, "school")
        for th in (1, 3):
            O.set_threads(th)
            assert np.array_equal(O.polmat_mul(A, B, p, "ntt"), ref), (m, k, n, la, lb, th)
    O.set_threads(0)


def test_ntt_extreme_values():
    """All residues p-1: the largest 128-bit sums the pointwise loop folds."""
    for p in (1152921504556515329, 4611686018429485057, 998244353):
        A = np.full((2, 70, 30), p - 1, dtype=np.uint64)
        B = np.full((70, 3, 20), p - 1, dtype=np.uint64)
        assert np.array_equal(O.polmat_mul(A, B, p, "ntt"), O.polmat_mul(A, B, p, "school"))


def test_ntt_rejects_short_2adic():
    p = 786433  # 3 * 2^18 + 1
    A = np.ones((1, 1, 2**18), dtype=np.uint64)
    with pytest.raises(ValueError):
        O.polmat_mul(A, A, p, "ntt")


def test_evaluation_identity():
    rng = np.random.default_rng(5)
    p = 1152921504556515329
    A, B = rand(rng, (5, 6, 300), p), rand(rng, (6, 4, 250), p)
    C = O.polmat_mul(A, B, p)
    for r in (0, 1, 2, 987654321, p - 1):
        Ar, Br = O.polmat_eval(A, r, p), O.polmat_eval(B, r, p)
        AB = np.array([[sum(int(Ar[i, t]) * int(Br[t, j]) for t in range(6)) % p for j in range(4)]
                       for i in range(5)], dtype=np.uint64)
        assert np.array_equal(O.polmat_eval(C, r, p), AB)


def test_mulmid_definition():
    rng = np.random.default_rng(6)
    for p in (3, 998244353, 1152921504556515329, 18446744073709551557):
        for _ in range(30):
            m, k, n = rng.integers(1, 5, size=3)
            la, lb = rng.integers(0, 40, size=2)
            nlo, nhi = int(rng.integers(0, 50)), int(rng.integers(0, 100))
            A, B = rand(rng, (m, k, la), p), rand(rng, (k, n, lb), p)
            full = O.polmat_mul(A, B, p)
            w = max(nhi - nlo, 0)
            ref = np.zeros((m, n, w), dtype=np.uint64)
            for d in range(w):
                if nlo + d < full.shape[2]:
                    ref[:, :, d] = full[:, :, nlo + d]
            assert np.array_equal(O.polmat_mulmid(A, B, nlo, nhi, p), ref)


# ---------------------------------------------------------------- rref / nullspace
def is_rref(R, rank, piv):
    r, c = R.shape
    for i in range(rank):
        row = R[i]
        nz = np.nonzero(row)[0]
        if len(nz) == 0 or nz[0] != piv[i] or row[piv[i]] != 1:
            return False
        if np.count_nonzero(R[:, piv[i]]) != 1:
            return False
        if i and piv[i] <= piv[i - 1]:
            return False
    return not R[rank:].any()


def test_rref_properties():
    rng = np.random.default_rng(7)
    for p in (2, 3, 7, 998244353, 1152921504556515329):
        for (r, c, rk) in [(5, 5, 5), (6, 9, 3), (9, 4, 2), (1, 1, 1), (4, 4, 0), (30, 60, 17)]:
            M = O.polmat_mul(rand(rng, (r, rk, 1), p), rand(rng, (rk, c, 1), p), p)[:, :, 0] \
                if rk else np.zeros((r, c), dtype=np.uint64)
            R, piv = O.rref(M, p)
            rank = len(piv)
            assert rank == O.rank(M, p) <= rk
            assert is_rref(R, rank, piv)
            # same row space: stacking adds no rank
            assert O.rank(np.vstack([M, R]), p) == rank


def nullspace_full(X, perm, m):
    nullity, rank = X.shape[0], m - X.shape[0]
    K = np.zeros((nullity, m), dtype=np.uint64)
    for i in range(nullity):
        K[i, perm[rank + i]] = 1
        for j in range(rank):
            K[i, perm[j]] = X[i, j]
    return K


def test_left_nullspace_convention():
    rng = np.random.default_rng(8)
    for p in (2, 5, 998244353, 1152921504556515329):
        for (m, n, rk) in [(6, 3, 3), (6, 3, 2), (8, 8, 8), (5, 7, 0), (7, 2, 1), (64, 32, 32), (10, 4, 4)]:
            A = O.polmat_mul(rand(rng, (m, rk, 1), p), rand(rng, (rk, n, 1), p), p)[:, :, 0] \
                if rk else np.zeros((m, n), dtype=np.uint64)
            rank = O.rank(A, p)
            nullity, X, perm = O.left_nullspace_compact(A, p)
            assert nullity == m - rank
            assert sorted(perm) == list(range(m))
            if nullity == 0:
                assert perm == list(range(m))
                continue
            # first part increasing, second part increasing
            assert perm[:rank] == sorted(perm[:rank]) and perm[rank:] == sorted(perm[rank:])
            K = nullspace_full(X, perm, m)
            # K A = 0 and K has full row rank nullity
            KA = O.polmat_mul(K[:, :, None], A[:, :, None], p)[:, :, 0]
            assert not KA.any()
            assert O.rank(K, p) == nullity
            # reduced echelon form with rightmost pivots: pivot of row i is its
            # last nonzero entry, pivots increase, pivot columns are unit vectors
            for i in range(nullity):
                nz = np.nonzero(K[i])[0]
                assert nz[-1] == perm[rank + i]
                assert np.count_nonzero(K[:, perm[rank + i]]) == 1
            # row rank profile: the first `rank` entries are independent rows of A
            assert O.rank(A[perm[:rank]], p) == rank


# ---------------------------------------------------------------- certificate
def test_certificate_rejects_planted_mistakes():
    rng = np.random.default_rng(9)
    p = 998244353
    m, n, order = 6, 3, 40
    F = rand(rng, (m, n, order), p)
    shift = np.zeros(m, dtype=np.int64)
    P, _ = O.pmbasis(F, order, shift, p)
    assert O.is_approximant_basis(P, F, order, shift, p) == 0
    # a coefficient changed -> residual no longer vanishes
    Q = P.copy()
    Q[1, 2, 0] = (int(Q[1, 2, 0]) + 1) % p
    assert O.is_approximant_basis(Q, F, order, shift, p) & 2
    # rows swapped -> not ordered weak Popov
    Q = P[[1, 0, 2, 3, 4, 5]].copy()
    assert O.is_approximant_basis(Q, F, order, shift, p) & 1
    # a row multiplied by x -> still approximants but no longer a basis
    Q = np.concatenate([P, np.zeros((m, m, 1), dtype=np.uint64)], axis=2)
    Q[0, :, 1:] = Q[0, :, :-1].copy()
    Q[0, :, 0] = 0
    assert O.is_approximant_basis(Q, F, order, shift, p) & (4 | 8)
    # a row scaled by a constant keeps everything valid
    Q = P.copy()
    Q[3] = (Q[3].astype(object) * 5 % p).astype(np.uint64)
    assert O.is_approximant_basis(Q, F, order, shift, p) == 0


def _pmul(a, b, p):
    if not a or not b:
        return []
    c = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            c[i + j] = (c[i + j] + x * y) % p
    return c


def _padd(a, b, p, sign=1):
    c = [0] * max(len(a), len(b))
    for i, x in enumerate(a):
        c[i] = x
    for i, y in enumerate(b):
        c[i] = (c[i] + sign * y) % p
    while c and c[-1] == 0:
        c.pop()
    return c


def _det_degree(P, p):
    """deg det(P) of a small square polynomial matrix (Leibniz formula over
    Python integers, -1 for det = 0): independent of the oracle."""
    m = P.shape[0]
    ent = [[[int(v) for v in P[i, j]] for j in range(m)] for i in range(m)]
    det = []
    for perm in itertools.permutations(range(m)):
        sgn = 1
        for i in range(m):
            for j in range(i + 1, m):
                if perm[i] > perm[j]:
                    sgn = -sgn
        term = [1]
        for i in range(m):
            term = _pmul(term, ent[i][perm[i]], p)
        det = _padd(det, term, p, sgn)
    return len(det) - 1


def _deg(v):
    nz = np.nonzero(v)[0]
    return int(nz[-1]) if len(nz) else -1


def _is_ord_weak_popov(P, shift):
    """pivot (rightmost entry reaching the shifted row degree) strictly
    increasing along the rows, no zero row"""
    prev = -1
    for i in range(P.shape[0]):
        piv, mx = -1, 0
        for j in range(P.shape[1]):
            d = _deg(P[i, j])
            if d >= 0 and (piv == -1 or mx <= d + shift[j]):
                piv, mx = j, d + shift[j]
        if piv == -1 or piv <= prev:
            return False
        prev = piv
    return True


def test_certificate_determinant_degree_is_the_references_test():
    """Bit 4 is set exactly when deg det(P) differs from the sum D of the
    diagonal degrees (verification.c:73-84), with deg det from an independent
    Leibniz expansion.  Random 3x3 and 4x4 matrices against a zero F (so only
    the form and determinant tests can fail); for the weak Popov ones the
    expansion also confirms deg det = D, the fact the oracle uses for them."""
    rng = np.random.default_rng(21)
    p = 998244353
    seen = {True: 0, False: 0}
    popov_seen = 0
    for trial in range(120):
        m = 3 + trial % 2
        L = 6
        P = rand(rng, (m, m, L), p)
        for i in range(m):
            for j in range(m):
                d = int(rng.integers(-1, L))
                P[i, j, d + 1:] = 0
                if d >= 0 and P[i, j, d] == 0:
                    P[i, j, d] = 1
        if trial % 3 == 0:  # make some of them ordered weak Popov (diagonal pivots)
            for i in range(m):
                di = int(rng.integers(0, L))
                P[i, i, :] = 0
                P[i, i, :di + 1] = rand(rng, di + 1, p)
                P[i, i, di] = 1 + P[i, i, di] % (p - 1)
                for j in range(m):
                    if j > i:
                        P[i, j, di:] = 0
                    elif j < i:
                        P[i, j, di + 1:] = 0
        shift = np.zeros(m, dtype=np.int64)
        F = np.zeros((m, 2, 3), dtype=np.uint64)
        D = sum(_deg(P[i, i]) for i in range(m))
        dd = _det_degree(P, p)
        popov = _is_ord_weak_popov(P, shift)
        got = O.is_approximant_basis(P, F, 2, shift, p)
        if popov:
            popov_seen += 1
            assert dd == D
            assert not got & 4
        else:
            assert bool(got & 4) == (dd != D), (trial, dd, D)
            seen[dd != D] += 1
    assert popov_seen >= 20 and seen[True] >= 10 and seen[False] >= 5, (popov_seen, seen)


def test_certificate_accepts_what_the_references_accepts():
    """Multiplying row 0 of a basis by (x + 1) keeps approximants, the weak
    Popov form, deg det = D and the rank of [P(0) | C]: the reference's
    certificate (a degree test, not a monomial test) accepts it, and so does
    this one; multiplying by x loses the rank of [P(0) | C] and is rejected."""
    rng = np.random.default_rng(4)
    p = 998244353
    m, n, order = 5, 2, 24
    F = rand(rng, (m, n, order), p)
    shift = np.zeros(m, dtype=np.int64)
    P, _ = O.pmbasis(F, order, shift, p)
    Q = np.concatenate([P, np.zeros((m, m, 1), dtype=np.uint64)], axis=2).astype(object)
    Q[0, :, 1:] = (Q[0, :, 1:] + Q[0, :, :-1]) % p
    assert O.is_approximant_basis(Q.astype(np.uint64), F, order, shift, p) == 0
    X = np.concatenate([P, np.zeros((m, m, 1), dtype=np.uint64)], axis=2)
    X[0, :, 1:] = X[0, :, :-1].copy()
    X[0, :, 0] = 0
    assert O.is_approximant_basis(X, F, order, shift, p) == 8


SHIFTS = {
    "uniform": lambda m, d: np.zeros(m),
    "increasing": lambda m, d: np.arange(m),
    "decreasing": lambda m, d: m - np.arange(m),
    "hermite": lambda m, d: np.arange(m) * m * d,
    "rhermite": lambda m, d: (m - np.arange(m)) * m * d,
    "plateau": lambda m, d: np.where(np.arange(m) < m // 2, 0, m * d),
    "rplateau": lambda m, d: np.where(np.arange(m) < m // 2, m * d, 0),
}


@pytest.mark.parametrize("shift_name", sorted(SHIFTS))
def test_pmbasis_certificate_collection(shift_name):
    """The shape of t-pmbasis.c:70-109 (random primes, dims, orders, shifts)."""
    rng = np.random.default_rng(10)
    primes = [p for p in PRIMES if p < 2**63]
    for it in range(12):
        p = primes[it % len(primes)]
        m, n = int(rng.integers(1, 9)), int(rng.integers(1, 7))
        order, ln = int(rng.integers(0, 90)), int(rng.integers(0, 90))
        F = rand(rng, (m, n, ln), p)
        shift = SHIFTS[shift_name](m, max(order, 1)).astype(np.int64)
        P, out_shift = O.pmbasis(F, order, shift, p)
        assert O.is_approximant_basis(P, F, order, shift, p) == 0
        # the returned shift is the shift-shifted row degree of P
        for i in range(m):
            degs = [int(np.nonzero(P[i, j])[0][-1]) + int(shift[j]) for j in range(m) if P[i, j].any()]
            assert out_shift[i] == max(degs)


def test_mbasis_order_one_closed_form():
    rng = np.random.default_rng(11)
    p = 1152921504556515329
    for (m, n) in [(5, 2), (4, 4), (7, 3)]:
        F = rand(rng, (m, n, 3), p)
        P, sh = O.mbasis(F, 1, np.zeros(m), p)
        nullity, X, perm = O.left_nullspace_compact(F[:, :, 0], p)
        rank = m - nullity
        ref = np.zeros((m, m, 2), dtype=np.uint64)
        for j in range(rank):
            ref[perm[j], perm[j], 1] = 1
        K = nullspace_full(X, perm, m)
        for t in range(nullity):
            ref[perm[rank + t], :, 0] = K[t]
        assert np.array_equal(O.normalise(P), O.normalise(ref))
        assert list(sh) == [1 if i in perm[:rank] else 0 for i in range(m)]


def test_basis_zero_input_is_identity():
    p = 998244353
    for fn in (O.mbasis, O.pmbasis):
        P, sh = fn(np.zeros((4, 2, 10), dtype=np.uint64), 50, np.arange(4), p)
        assert np.array_equal(O.normalise(P)[:, :, 0], np.eye(4, dtype=np.uint64))
        assert list(sh) == [0, 1, 2, 3]


def test_mbasis_equals_pmbasis_generic():
    rng = np.random.default_rng(12)
    for p in (998244353, 1152921504556515329):
        for (m, n, order) in [(4, 2, 70), (6, 3, 100), (8, 4, 64)]:
            F = rand(rng, (m, n, order), p)
            sh = rng.integers(0, 5, size=m)
            P1, s1 = O.mbasis(F, order, sh, p)
            P2, s2 = O.pmbasis(F, order, sh, p)
            assert np.array_equal(O.normalise(P1), O.normalise(P2))
            assert np.array_equal(s1, s2)


def test_mbasis_equals_pmbasis_degenerate():
    """The divide and conquer repeats M-Basis's steps exactly (each step
    depends only on the residual coefficient and the shifts, and the updates
    are row operations, which commute with the product by the first half's
    basis), so the basis does not depend on where the recursion stops: the
    iterative basis equals the recursive one also where pivots are scarce
    (primes 2 to 7, mostly zero or repeated rows, zero columns, wide shifts)."""
    rng = np.random.default_rng(1000)
    for it in range(60):
        p = (2, 3, 5, 7, 101, 998244353)[it % 6]
        m, n = int(rng.integers(1, 9)), int(rng.integers(1, 9))
        order = int(rng.integers(33, 130))
        lenF = int(rng.integers(1, order + 10))
        F = rng.integers(0, p, size=(m, n, lenF), dtype=np.uint64)
        kind = it % 5
        if kind == 1:
            F[:, :, rng.random(lenF) < 0.7] = 0
        if kind == 2 and m > 1:
            F[m - 1] = F[0]
        if kind == 3:
            F[:, 0] = 0
        if kind == 4:
            F[rng.random((m, n, lenF)) < 0.9] = 0
        sh = rng.integers(-5, 40, size=m).astype(np.int64)
        P1, s1 = O.mbasis(F, order, sh, p)
        P2, s2 = O.pmbasis(F, order, sh, p)
        assert np.array_equal(O.normalise(P1), O.normalise(P2)), it
        assert np.array_equal(s1, s2), it


def test_golden_fixture_small_entries():
    """The committed fixture agrees with the oracle on the entries that are
    cheap to recompute (the others are checked on the GPU at full size)."""
    import hashlib
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))
    e = g["c1_mul_4x4x4_len256_p60"]
    A = O.rand_uniform((4, 4, 256), e["p"], e["seedA"])
    B = O.rand_uniform((4, 4, 256), e["p"], e["seedB"])
    C = O.polmat_mul(A, B, e["p"], "ntt")
    assert hashlib.sha256(C.tobytes()).hexdigest() == e["sha256"]
    assert g["c3_pmbasis_64x32_order2048_p30"]["certificate"] == 0


def test_rand_uniform_stream():
    """SplitMix64 stream (fixed constants) and rejection to [0, p)."""
    x = O.rand_uniform((3,), 2**61, 7)
    st, vals = 7, []
    for _ in range(3):
        st = (st + 0x9E3779B97F4A7C15) % 2**64
        z = st
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) % 2**64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) % 2**64
        vals.append((z ^ (z >> 31)) >> 3)
    assert [int(v) for v in x] == vals
    y = O.rand_uniform((1000,), 998244353, 1)
    assert int(y.max()) < 998244353
