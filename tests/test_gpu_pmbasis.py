"""GPU parity of approximant bases (PM-Basis, M-Basis) with the oracle.

Follows the shape of the reference tests (flint-extras/src/nmod_poly_mat_extra/
test/t-pmbasis.c:70-109 and testing_collection.h: random primes, dimensions,
orders, the eight shift families, zero and rank-deficient inputs), but checks
bit-exact equality with the oracle's restatement of the reference algorithm,
on top of the certificate the reference uses.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

P60 = 1152921504556515329
P30 = 998244353
PRIMES_ANY = [2, 3, 5, 521, 524309, 536870923, 549755813911, 576460752303423619]  # order <= 32: leaf only
PRIMES_FFT = [P30, P60, 7340033, 469762049]


@pytest.fixture(scope="module")
def pml():
    import pml_amd
    return pml_amd


_ctxs = {}


def ctx_for(pml, p):
    if p not in _ctxs:
        _ctxs[p] = pml.Context(p)
    return _ctxs[p]


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


def make(rng, kind, m, n, ln, p):
    if kind == "uniform":
        return rng.integers(0, p, (m, n, ln), dtype=np.uint64)
    if kind == "zero":
        return np.zeros((m, n, ln), dtype=np.uint64)
    if kind == "sparse":
        F = rng.integers(0, p, (m, n, ln), dtype=np.uint64)
        return F * (rng.random((m, n, ln)) < 0.05)
    if kind == "rkdef":  # rank at most 1 at every coefficient
        u = rng.integers(0, p, (m, 1, ln), dtype=np.uint64)
        v = rng.integers(0, min(p, 4), (1, n, 1), dtype=np.uint64)
        return (u.astype(object) * v.astype(object) % p).astype(np.uint64)
    if kind == "cdeg":  # unbalanced column degrees
        F = rng.integers(0, p, (m, n, ln), dtype=np.uint64)
        for j in range(n):
            F[:, j, max(1, ln * (j + 1) // (n + 1)):] = 0
        return F


def check(pml, p, F, order, shift, fn="pmbasis"):
    P_ref, s_ref = getattr(O, fn)(F, order, shift, p)
    P_dev, s_dev = getattr(ctx_for(pml, p), fn)(F, order, shift)
    assert np.array_equal(O.normalise(P_dev), O.normalise(P_ref))
    assert np.array_equal(s_dev, s_ref)
    return P_dev


@pytest.mark.parametrize("shift_name", sorted(SHIFTS))
def test_pmbasis_shift_families(pml, shift_name):
    rng = np.random.default_rng(31)
    for it in range(6):
        p = PRIMES_FFT[it % len(PRIMES_FFT)]
        m, n = int(rng.integers(1, 10)), int(rng.integers(1, 8))
        order = int(rng.integers(1, 150))
        F = make(rng, "uniform", m, n, int(rng.integers(0, 150)), p)
        shift = SHIFTS[shift_name](m, order).astype(np.int64)
        P = check(pml, p, F, order, shift)
        assert O.is_approximant_basis(P, F, order, shift, p) == 0


@pytest.mark.parametrize("kind", ["zero", "sparse", "rkdef", "cdeg"])
def test_pmbasis_matrix_families(pml, kind):
    rng = np.random.default_rng(37)
    for it in range(4):
        p = PRIMES_FFT[it % len(PRIMES_FFT)]
        m, n = int(rng.integers(2, 9)), int(rng.integers(1, 6))
        order = int(rng.integers(1, 120))
        F = make(rng, kind, m, n, order, p)
        check(pml, p, F, order, rng.integers(0, 4, m).astype(np.int64))


@pytest.mark.parametrize("p", PRIMES_ANY)
def test_leaf_any_prime(pml, p):
    """Orders up to 32 run only the device M-Basis step (no transform), so
    every word-size prime below 2^62 is allowed, as in the reference tests."""
    rng = np.random.default_rng(p % 9973)
    for (m, n, order) in [(4, 2, 32), (5, 5, 20), (3, 1, 31), (9, 4, 17)]:
        F = make(rng, "uniform", m, n, order + 3, p)
        sh = rng.integers(0, 3, m).astype(np.int64)
        check(pml, p, F, order, sh, "pmbasis")
        check(pml, p, F, order, sh, "mbasis")


def test_pmbasis_prime_two(pml):
    """p = 2 (the reference's tests can draw it: a random 2-bit prime):
    plain-residue leaf arithmetic, products through the multi-prime path."""
    rng = np.random.default_rng(2)
    for (m, n, order, kind) in [(4, 2, 100, "uniform"), (7, 3, 200, "sparse"), (3, 1, 65, "uniform"),
                                (6, 4, 130, "rkdef")]:
        F = make(rng, kind, m, n, order, 2)
        sh = rng.integers(0, 3, m).astype(np.int64)
        P = check(pml, 2, F, order, sh)
        assert O.is_approximant_basis(P, F, order, sh, 2) == 0


def test_mbasis_long_order(pml):
    rng = np.random.default_rng(41)
    for p in (P30, P60):
        F = make(rng, "uniform", 6, 2, 90, p)
        check(pml, p, F, 90, np.zeros(6, dtype=np.int64), "mbasis")


def test_full_rank_residual_stops(pml):
    """n >= m: the residual reaches full row rank, so the basis is X^k times
    the current one and the loop stops (nmod_mat_poly_mbasis.c:221-231)."""
    rng = np.random.default_rng(43)
    for p in (P30, P60):
        for (m, n, order) in [(3, 5, 40), (4, 4, 70), (1, 3, 50)]:
            F = make(rng, "uniform", m, n, order, p)
            check(pml, p, F, order, np.zeros(m, dtype=np.int64))


def test_pmbasis_dims_64(pml):
    rng = np.random.default_rng(47)
    F = make(rng, "uniform", 64, 32, 200, P30)
    P = check(pml, P30, F, 200, np.zeros(64, dtype=np.int64))
    assert O.is_approximant_basis(P, F, 200, np.zeros(64, dtype=np.int64), P30) == 0


@pytest.mark.parametrize("m,n,order,p,kind,shift_name", [
    (100, 70, 40, P30, "uniform", "increasing"),
    (128, 100, 36, P60, "uniform", "uniform"),
    (72, 65, 70, P30, "rkdef", "plateau"),
    (128, 128, 3, P60, "uniform", "shuffle"),
    (96, 90, 33, P60, "sparse", "decreasing"),
])
def test_pmbasis_wide_residual(pml, m, n, order, p, kind, shift_name):
    # more than 64 residual columns: the composing wave holds two rows of the
    # transposed residual per lane
    rng = np.random.default_rng(m * 1000 + n)
    F = make(rng, kind, m, n, order, p)
    shift = SHIFTS[shift_name](m, order).astype(np.int64)
    P = check(pml, p, F, order, shift)
    assert O.is_approximant_basis(P, F, order, shift, p) == 0


@pytest.mark.parametrize("m,n,lenF,order,p", [
    (6, 3, 40, 40, P30),      # one leaf level below the root
    (9, 4, 150, 100, P60),    # F longer than the order
    (5, 5, 20, 70, P30),      # F shorter than the order
    (64, 32, 200, 200, P60),
])
def test_pmbasis_device_resident(pml, m, n, lenF, order, p):
    # pmlg_pmbasis_dev on device buffers gives the host entry point's basis
    rng = np.random.default_rng(m * 7 + order)
    F = rng.integers(0, p, (m, n, lenF), dtype=np.uint64)
    shift = rng.integers(-5, 20, m).astype(np.int64)
    ctx = ctx_for(pml, p)
    P, sh = ctx.pmbasis(F, order, shift)
    cap = order + 1
    dF = pml.DeviceArray(ctx, F.nbytes)
    dF.upload(F)
    dP = pml.DeviceArray(ctx, m * m * cap * 8)
    dP.upload(np.full((m, m, cap), np.uint64(2**64 - 1), dtype=np.uint64))
    sh_d = shift.copy()
    L = ctx.pmbasis_dev(dP.ptr, cap, sh_d, dF.ptr, m, n, lenF, order)
    # the device length is the computed one: the host entry point's basis
    # followed by zero coefficients
    Lh = P.shape[2]
    assert Lh <= L <= cap
    D = dP.download(np.empty((m, m, cap), dtype=np.uint64))[:, :, :L]
    assert np.array_equal(D[:, :, :Lh], P)
    assert not D[:, :, Lh:].any()
    assert np.array_equal(sh_d, sh)


SWEEP_DIMS = [(1, 1), (2, 1), (31, 17), (32, 32), (33, 20), (40, 33), (63, 63), (64, 64),
              (65, 40), (66, 65), (96, 70), (127, 100), (128, 1), (128, 64), (128, 65),
              (128, 128), (5, 100), (17, 128), (100, 32), (128, 31), (256, 32)]


@pytest.mark.parametrize("p", [P30, P60])
@pytest.mark.parametrize("m,n", SWEEP_DIMS)
def test_pmbasis_dimension_sweep(pml, m, n, p):
    # the step kernel's column blocks change at m = 32 and m = 64, its rows per
    # lane at n = 64, its composition (coefficients on the wave's upper half)
    # at n = 32, its panel width with the word size; short orders keep the
    # oracle quick, and include residuals of full and deficient rank
    rng = np.random.default_rng(m * 1009 + n * 7 + (p & 0xff))
    order = int(rng.integers(3, 9))
    F = make(rng, "uniform" if (m + n) % 2 else "rkdef", m, n, order, p)
    shift = rng.integers(-3, 6, m).astype(np.int64)
    check(pml, p, F, order, shift, fn="mbasis")
    P = check(pml, p, F, order, shift)
    assert O.is_approximant_basis(P, F, order, shift, p) == 0


@pytest.mark.parametrize("p", [P30, P60, 674830283164631843])
def test_pmbasis_leaf_order_does_not_change_basis(pml, p):
    """The M-Basis / PM-Basis crossover (PMBASIS_THRES,
    nmod_poly_mat_approximant.h:101) is a speed choice only: whatever the leaf
    order, the basis and shifts equal the oracle's (which splits at 32, as the
    reference does), on generic and degenerate inputs; the non-FFT prime runs
    the products through the multi-prime path."""
    ctx = ctx_for(pml, p)
    rng = np.random.default_rng(p % 7919)
    try:
        for kind, (m, n, order) in [("uniform", (6, 3, 300)), ("rkdef", (5, 2, 257)),
                                    ("sparse", (4, 3, 200)), ("cdeg", (7, 4, 130))]:
            F = make(rng, kind, m, n, order, p)
            shift = rng.integers(-2, 5, m).astype(np.int64)
            P_ref, s_ref = O.pmbasis(F, order, shift, p)
            for t in (1, 7, 16, 64, 128, 256):
                ctx.set_leaf_order(t)
                P_dev, s_dev = ctx.pmbasis(F, order, shift)
                assert np.array_equal(O.normalise(P_dev), O.normalise(P_ref)), (kind, t)
                assert np.array_equal(s_dev, s_ref), (kind, t)
    finally:
        ctx.set_leaf_order(0)


def test_pmbasis_measured_leaf_order(pml):
    """Calls of order >= 512 use the leaf order measured for their shape; the
    basis is still the oracle's."""
    ctx = ctx_for(pml, P30)
    ctx.set_leaf_order(0)
    t = ctx.leaf_order(8, 4)
    assert t in (16, 32, 64, 128, 256)
    assert ctx.leaf_order(8, 4) == t  # kept per shape
    rng = np.random.default_rng(53)
    F = make(rng, "uniform", 8, 4, 600, P30)
    shift = rng.integers(0, 3, 8).astype(np.int64)
    P = check(pml, P30, F, 600, shift)
    assert O.is_approximant_basis(P, F, 600, shift, P30) == 0
    for bad in (-1, 257):
        with pytest.raises(pml.PmlgError):
            ctx.set_leaf_order(bad)


# Bases of more than 128 rows (up to 256; at most 128 columns, within one
# workgroup's LDS: n (m | 1) words for the transposed residual).  The
# reference has no such cap (approximant_basis.c:33-69,
# nmod_mat_poly_mbasis.c:116-378).
WIDE = [
    # m, n, order, p, iterative
    (160, 80, 40, P60, False),
    (200, 50, 48, P30, False),
    (256, 62, 24, P60, False),   # the largest n for m = 256 with 64-bit words
    (256, 100, 20, P30, False),  # 32-bit words: n up to 128 at m = 256
    (150, 60, 20, P60, True),
    (129, 128, 36, P30, False),
]


@pytest.mark.parametrize("case", WIDE)
def test_bases_with_more_than_128_rows(pml, case):
    m, n, order, p, iterative = case
    F = O.rand_uniform((m, n, order), p, m + n)
    shift = np.random.default_rng(m * n).integers(0, 4, size=m).astype(np.int64)
    ref_fn = O.mbasis if iterative else O.pmbasis
    P_ref, s_ref = ref_fn(F, order, shift, p)
    ctx = ctx_for(pml, p)
    P_dev, s_dev = (ctx.mbasis(F, order, shift) if iterative else ctx.pmbasis(F, order, shift))
    assert np.array_equal(O.normalise(P_dev), O.normalise(P_ref)), case
    assert np.array_equal(s_dev, s_ref)
    assert O.is_approximant_basis(O.normalise(P_dev), F, order, shift, p) == 0
    print("basis %s identical" % (case,), flush=True)


def test_wide_basis_beyond_lds_refused(pml):
    ctx = ctx_for(pml, P60)
    F = O.rand_uniform((256, 64, 8), P60, 1)
    with pytest.raises(pml.PmlgError) as ei:
        ctx.pmbasis(F, 8, np.zeros(256, dtype=np.int64))
    assert ei.value.status == 2 and "LDS" in str(ei.value)
    with pytest.raises(pml.PmlgError) as ei:
        ctx.pmbasis(O.rand_uniform((257, 4, 8), P60, 1), 8, np.zeros(257, dtype=np.int64))
    assert ei.value.status == 1


@pytest.mark.parametrize("seed", [11, 12])
def test_pmbasis_randomized_sweep(pml, seed):
    # Randomized like the reference's t-pmbasis.c:70-109 (random prime,
    # dimensions, order and shifts per iteration): 150 seeded draws over
    # FFT and non-FFT primes (p = 2 and a 59-bit prime included), matrix
    # families and shift families, M-Basis and PM-Basis, each identical to
    # the oracle's basis and shift
    rng = np.random.default_rng(seed)
    primes = [P30, P60, 97, 7340033, 3, 576460752303423619, 2]
    small = (2, 3, 97, 576460752303423619)
    for it in range(150):
        p = primes[rng.integers(len(primes))]
        m = int(rng.integers(1, 90))
        n = int(rng.integers(1, 40 if rng.random() < 0.7 else 100))
        order = int(rng.integers(1, 24 if p in small else 48))
        kind = ["uniform", "sparse", "rkdef", "cdeg", "zero"][rng.integers(5)]
        F = make(rng, kind, m, n, order, p)
        sname = sorted(SHIFTS)[rng.integers(len(SHIFTS))]
        shift = SHIFTS[sname](m, order).astype(np.int64) + rng.integers(-2, 3, m)
        fn = "mbasis" if rng.random() < 0.3 else "pmbasis"
        P_ref, s_ref = getattr(O, fn)(F, order, shift, p)
        P_dev, s_dev = getattr(ctx_for(pml, p), fn)(F, order, shift)
        case = (it, fn, p, m, n, order, kind, sname)
        assert np.array_equal(O.normalise(P_dev), O.normalise(P_ref)), case
        assert np.array_equal(s_dev, s_ref), case
