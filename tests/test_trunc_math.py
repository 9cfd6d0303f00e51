"""The algebra of the 3/4-truncated products, checked on the CPU.

A product whose length lenC satisfies N/2 < lenC <= 3N/4 is evaluated at the
frequencies f with f mod 4 in {0, 1, 2} only.  The device recovers the
coefficients in one of two ways (pml_amd/csrc/ntt.hip):

* unsplit (trunc_correct): the N-point inverse runs with the missing quarter
  set to zero, giving y; with M = N/4 and s = w^M the coefficients are
  c_q = y_q - s^(q+1) y_3 at positions r + q M, q = 0, 1, 2;
* split in four sub-transforms of M points (trunc_solve): with a, v, u the
  inverses of sub-transforms 0, 1, 2 at position r, twisted by w^(-g r), the
  outputs are x1 = (2/N)(a - u), x0 = S + D - (s/2) x1, x2 = S - D + (s/2) x1,
  where S = (a + u)/N and D = (2/N) v.

Both are checked here against the schoolbook product, with exact DFTs over
Z/pZ written out in full, for several primes, lengths and random operands.
"""
import numpy as np
import pytest

PRIMES = [998244353, 7340033, 469762049]


def root_of_unity(p, n):
    """An element of exact order n (n a power of two dividing p - 1)."""
    assert (p - 1) % n == 0
    for g in range(2, p):
        w = pow(g, (p - 1) // n, p)
        if pow(w, n // 2, p) != 1:
            return w
    raise AssertionError("no root")


def schoolbook(a, b, p):
    c = [0] * (len(a) + len(b) - 1)
    for i, x in enumerate(a):
        for j, y in enumerate(b):
            c[i + j] = (c[i + j] + x * y) % p
    return c


def dft(coeffs, w, n, p):
    return [sum(c * pow(w, i * f, p) for i, c in enumerate(coeffs)) % p for f in range(n)]


def idft_unscaled(vals, w, n, p):
    wi = pow(w, p - 2, p)
    return [sum(v * pow(wi, i * f, p) for f, v in enumerate(vals)) % p for i in range(n)]


def truncated_unsplit(A, B, w, N, p):
    M = N // 4
    s = pow(w, M, p)
    X = [x * y % p for x, y in zip(A, B)]
    for f in range(N):
        if f % 4 == 3:
            X[f] = 0  # the dropped evaluations
    ninv = pow(N, p - 2, p)
    y = [v * ninv % p for v in idft_unscaled(X, w, N, p)]
    c = [0] * N
    for r in range(M):
        y3 = y[r + 3 * M]
        for q in range(3):
            c[r + q * M] = (y[r + q * M] - pow(s, q + 1, p) * y3) % p
    return c


def truncated_split(A, B, w, N, p):
    M = N // 4
    s = pow(w, M, p)
    w4 = pow(w, 4, p)
    z = []
    for g in range(3):
        sub = [A[4 * h + g] * B[4 * h + g] % p for h in range(M)]
        z.append(idft_unscaled(sub, w4, M, p))  # M w^(r g) sum_q c_(r+qM) s^(q g)
    wi = pow(w, p - 2, p)
    inv = lambda x: pow(x % p, p - 2, p)
    c = [0] * N
    for r in range(M):
        a = z[0][r]
        v = z[1][r] * pow(wi, r, p) % p
        u = z[2][r] * pow(wi, 2 * r, p) % p
        x1 = 2 * inv(N) * (a - u) % p
        S = (a + u) * inv(N) % p
        D = 2 * inv(N) * v % p
        h = s * inv(2) % p * x1 % p
        c[r] = (S + D - h) % p
        c[r + M] = x1
        c[r + 2 * M] = (S - D + h) % p
    return c


@pytest.mark.parametrize("p", PRIMES)
@pytest.mark.parametrize("logN", [4, 5, 6])
@pytest.mark.parametrize("form", ["unsplit", "split"])
def test_truncated_products_equal_schoolbook(p, logN, form):
    N = 1 << logN
    w = root_of_unity(p, N)
    rng = np.random.default_rng(p % 1000 + logN)
    # product lengths across (N/2, 3N/4]: the range the device truncates
    for lenc in (N // 2 + 1, (5 * N) // 8, 3 * N // 4):
        la = int(rng.integers(1, lenc + 1))
        lb = lenc + 1 - la
        a = [int(x) for x in rng.integers(0, p, size=la)]
        b = [int(x) for x in rng.integers(0, p, size=lb)]
        A, B = dft(a, w, N, p), dft(b, w, N, p)
        got = (truncated_unsplit if form == "unsplit" else truncated_split)(A, B, w, N, p)
        ref = schoolbook(a, b, p)
        assert got[:lenc] == ref, (p, N, lenc, form)
        assert not any(got[lenc:3 * N // 4]), (p, N, lenc, form)


def test_truncation_needs_the_zero_last_quarter():
    """Past 3N/4 the correction is wrong: the recovered coefficients differ
    from the product (so the test above is not passing by accident)."""
    p, N = 998244353, 32
    w = root_of_unity(p, N)
    rng = np.random.default_rng(5)
    lenc = 3 * N // 4 + 3
    a = [int(x) for x in rng.integers(0, p, size=lenc // 2)]
    b = [int(x) for x in rng.integers(0, p, size=lenc + 1 - lenc // 2)]
    A, B = dft(a, w, N, p), dft(b, w, N, p)
    ref = schoolbook(a, b, p)
    assert truncated_unsplit(A, B, w, N, p)[:3 * N // 4] != ref[:3 * N // 4]
    assert truncated_split(A, B, w, N, p)[:3 * N // 4] != ref[:3 * N // 4]
