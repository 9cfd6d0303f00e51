"""GPU parity of the polynomial-matrix product and middle product.

The device path (libpmlgpu.so, through its C ABI) is compared bit for bit
with the CPU oracle on the same seeded inputs.  The reference tests the same
operations by randomized equality against FLINT (flint-extras/src/
nmod_poly_mat_extra/test/t-mul_geometric.c:24-80, t-mulmid.c:20-107); the
product being unique, equality with the oracle is the same check.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

P60 = 1152921504556515329   # 2^60 - 3*2^24 + 1, v2(p-1) = 24
P30 = 998244353             # 119*2^23 + 1
P62 = 4611686018326724609   # 61.99-bit prime with v2(p-1) >= 20 (plain kernel only)
SMALL_FFT = [786433, 7340033, 469762049]


@pytest.fixture(scope="module")
def pml():
    import pml_amd
    return pml_amd


_ctxs = {}


def ctx_for(pml, p):
    if p not in _ctxs:
        _ctxs[p] = pml.Context(p)
    return _ctxs[p]


def rand(rng, shape, p):
    return rng.integers(0, p, size=shape, dtype=np.uint64)


CASES = [
    # m, k, n, lenA, lenB
    (1, 1, 1, 1, 1),
    (2, 3, 4, 5, 7),
    (4, 4, 4, 256, 256),      # config C1 shape
    (3, 5, 2, 100, 33),
    (7, 33, 9, 64, 65),       # ragged, crosses a 32 tile
    (33, 2, 31, 17, 300),
    (40, 40, 40, 40, 40),
    (1, 64, 1, 1000, 1000),
    (5, 1, 6, 129, 1),
]


@pytest.mark.parametrize("p", [P60, P30] + SMALL_FFT)
@pytest.mark.parametrize("method", ["plain", "mfma"])
@pytest.mark.parametrize("case", CASES)
def test_mul_matches_oracle(pml, p, method, case):
    m, k, n, la, lb = case
    rng = np.random.default_rng(hash((p, case)) & 0xFFFF)
    A, B = rand(rng, (m, k, la), p), rand(rng, (k, n, lb), p)
    C = ctx_for(pml, p).mul(A, B, method=method)
    assert np.array_equal(C, O.polmat_mul(A, B, p))


@pytest.mark.parametrize("p,method", [(P60, "plain"), (P60, "mfma"), (P30, "plain"), (P30, "mfma"),
                                      (P62, "plain"), (7340033, "mfma")])
def test_mul_truncated_transforms(pml, p, method):
    """Products of length at most 3N/4 evaluate at the first 3N/4 positions
    of the transform only (the DIF outputs at the frequencies 0, 1, 2 mod 4)
    and rebuild the top quarter from the known zero coefficients: both
    boundaries (N/2 + 1 and 3N/4) and one length just above (a full
    transform), for the LDS kernels (N < 2^11), the register kernels and the
    split transforms."""
    rng = np.random.default_rng(p % 1000 + len(method))
    # up to 4 times the longest sub-transform (2^13 points of 64 bits, 2^14
    # of 32 bits): beyond one sub-transform the transform is split in four and
    # the fourth sub-transform dropped
    logs = [5, 6, 9, 10, 11, 12, 13, 14, 15] + ([16] if p < (1 << 30) else [])
    for lg in logs:
        N = 1 << lg
        for lenc in (N // 2 + 1, 3 * N // 4 - 1, 3 * N // 4, 3 * N // 4 + 1):
            la = lenc // 3 + 1
            lb = lenc - la + 1
            A, B = rand(rng, (2, 3, la), p), rand(rng, (3, 2, lb), p)
            C = ctx_for(pml, p).mul(A, B, method=method)
            assert np.array_equal(C, O.polmat_mul(A, B, p, method="ntt")), (lg, lenc)


@pytest.mark.parametrize("method", ["plain", "mfma"])
def test_mul_large_length_split(pml, method):
    """Transform lengths above the LDS limit (radix-2/4 split + merge)."""
    rng = np.random.default_rng(7)
    for p, la, lb in [(P60, 9000, 7000), (P60, 16384, 16384), (P30, 20000, 12000)]:
        A, B = rand(rng, (2, 2, la), p), rand(rng, (2, 1, lb), p)
        C = ctx_for(pml, p).mul(A, B, method=method)
        assert np.array_equal(C, O.polmat_mul(A, B, p, method="ntt"))


def test_mul_extreme_residues(pml):
    """All coefficients p-1 (largest limbs, largest accumulations)."""
    for p in (P60, P30):
        A = np.full((32, 32, 64), p - 1, dtype=np.uint64)
        B = np.full((32, 32, 64), p - 1, dtype=np.uint64)
        ref = O.polmat_mul(A, B, p)
        for method in ("plain", "mfma"):
            assert np.array_equal(ctx_for(pml, p).mul(A, B, method=method), ref)


def test_mul_plain_62bit(pml):
    rng = np.random.default_rng(3)
    A, B = rand(rng, (3, 4, 50), P62), rand(rng, (4, 5, 60), P62)
    ctx = ctx_for(pml, P62)
    assert np.array_equal(ctx.mul(A, B, method="plain"), O.polmat_mul(A, B, P62))
    assert np.array_equal(ctx.mul(A, B, method="auto"), O.polmat_mul(A, B, P62))
    with pytest.raises(pml.PmlgError):
        ctx.mul(A, B, method="mfma")


def test_mul_zero_and_empty(pml):
    ctx = ctx_for(pml, P60)
    A = np.zeros((3, 4, 10), dtype=np.uint64)
    B = np.zeros((4, 2, 5), dtype=np.uint64)
    assert not ctx.mul(A, B).any()
    assert ctx.mul(np.zeros((3, 4, 0), np.uint64), B).shape == (3, 2, 0)
    assert ctx.mul(np.zeros((3, 0, 3), np.uint64), np.zeros((0, 2, 3), np.uint64)).shape == (3, 2, 5)


def test_length_beyond_the_roots_takes_the_multiprime_path(pml):
    # 786433 - 1 = 3 * 2^18: a product of length 2^19 needs a root the field
    # lacks, so it is computed modulo the three CRT primes
    p = 786433
    ctx = ctx_for(pml, p)
    A = O.rand_uniform((1, 1, 300000), p, 4)
    C = ctx.mul(A, A)
    for r in (5, 777):
        a = int(O.polmat_eval(A, r, p)[0, 0])
        assert int(O.polmat_eval(C, r, p)[0, 0]) == a * a % p


MULMID_CASES = [
    # m, k, n, lenA, lenB, nlo, nhi
    (3, 4, 5, 10, 20, 5, 25),
    (8, 8, 8, 33, 64, 32, 64),      # pmbasis-like: len1 <= nlo+1 (wrapped transform)
    (8, 8, 8, 64, 33, 32, 64),
    (2, 3, 2, 50, 50, 0, 99),
    (2, 3, 2, 50, 50, 10, 200),     # nhi beyond the product
    (2, 3, 2, 50, 50, 120, 200),    # nlo beyond the product
    (4, 4, 4, 40, 40, 30, 30),      # empty range
    (64, 64, 32, 513, 1024, 512, 1024),
]


@pytest.mark.parametrize("p", [P60, P30])
@pytest.mark.parametrize("method", ["plain", "mfma"])
@pytest.mark.parametrize("case", MULMID_CASES)
def test_mulmid_matches_oracle(pml, p, method, case):
    m, k, n, la, lb, nlo, nhi = case
    rng = np.random.default_rng(11)
    A, B = rand(rng, (m, k, la), p), rand(rng, (k, n, lb), p)
    out = ctx_for(pml, p).mulmid(A, B, nlo, nhi, method=method)
    assert np.array_equal(out, O.polmat_mulmid(A, B, nlo, nhi, p))


def test_config2_shape_property(pml):
    """32x32 by 32x32, length 4096, 60-bit prime (the benchmark shape): both
    kernels agree and C(r) = A(r) B(r) at seeded random points."""
    rng = np.random.default_rng(2024)
    p = P60
    A, B = rand(rng, (32, 32, 4096), p), rand(rng, (32, 32, 4096), p)
    ctx = ctx_for(pml, p)
    C1 = ctx.mul(A, B, method="mfma")
    C2 = ctx.mul(A, B, method="plain")
    assert np.array_equal(C1, C2)
    for r in (3, 123456789, p - 2):
        Ar, Br, Cr = O.polmat_eval(A, r, p), O.polmat_eval(B, r, p), O.polmat_eval(C1, r, p)
        AB = O.polmat_mul(Ar[:, :, None], Br[:, :, None], p)[:, :, 0]
        assert np.array_equal(AB, Cr)
    # one entry in full against the closed form
    c00 = O.polmat_mul(A[:1], B[:, :1], p, method="ntt")
    assert np.array_equal(C1[:1, :1], c00)


def test_mfma_inner_dimension_bound(pml):
    """The int8-limb kernel rebuilds the exact dot product V in 128 bits and
    reduces it by Montgomery REDC, which needs V + (p+1) 2^64 <= 2^128, i.e.
    k <= 239 for this prime; beyond, 'auto' uses the plain kernel and 'mfma'
    refuses."""
    rng = np.random.default_rng(17)
    p = P60
    ctx = ctx_for(pml, p)
    A, B = rand(rng, (3, 239, 9), p), rand(rng, (239, 2, 7), p)
    assert np.array_equal(ctx.mul(A, B, method="mfma"), O.polmat_mul(A, B, p))
    # worst case for the accumulators: every coefficient p-1
    A, B = np.full((2, 239, 3), p - 1, np.uint64), np.full((239, 2, 3), p - 1, np.uint64)
    assert np.array_equal(ctx.mul(A, B, method="mfma"), O.polmat_mul(A, B, p))
    A, B = rand(rng, (2, 300, 5), p), rand(rng, (300, 2, 5), p)
    assert np.array_equal(ctx.mul(A, B, method="auto"), O.polmat_mul(A, B, p))
    with pytest.raises(pml.PmlgError):
        ctx.mul(A, B, method="mfma")


@pytest.mark.parametrize("seed", [21, 22])
def test_mul_randomized_sweep(pml, seed):
    # Randomized like the reference's t-mul_geometric.c:24-80 (a random
    # prime, dimensions and lengths per iteration): 200 seeded products or
    # middle products over FFT primes, non-FFT primes (the CRT path) and
    # moduli up to 2^64, with each form, identical to the oracle's
    rng = np.random.default_rng(seed)
    primes = [P30, P60, 7340033, 469762049, 97, 674830283164631843, 2**64 - 59, 2**63 + 29,
              1000000007, 2]
    for it in range(200):
        p = int(primes[rng.integers(len(primes))])
        fft = p in (P30, P60, 7340033, 469762049)
        m, k, n = (int(x) for x in rng.integers(1, 20 if fft else 7, 3))
        top = 3000 if fft else 200
        la, lb = int(rng.integers(1, top)), int(rng.integers(1, top))
        if rng.random() < 0.3:
            la = int(rng.integers(1, 20))
        A = O.rand_uniform((m, k, la), p, int(rng.integers(1 << 30)))
        B = O.rand_uniform((k, n, lb), p, int(rng.integers(1 << 30)))
        method = ["auto", "plain", "mfma"][rng.integers(3)]
        ctx = ctx_for(pml, p)
        case = (it, p, m, k, n, la, lb, method)
        if rng.random() < 0.7:
            assert np.array_equal(ctx.mul(A, B, method=method), O.polmat_mul(A, B, p)), case
        else:
            lc = la + lb - 1
            nlo = int(rng.integers(0, lc))
            nhi = int(rng.integers(nlo + 1, lc + 1))
            assert np.array_equal(ctx.mulmid(A, B, nlo, nhi, method=method),
                                  O.polmat_mulmid(A, B, nlo, nhi, p)), case
