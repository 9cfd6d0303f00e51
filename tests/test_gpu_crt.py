"""GPU parity of the multi-prime (CRT) path: products, middle products and
approximant bases modulo primes without the roots of unity a transform needs
(the reference's own benchmark primes, ntl-extras/benchmarks/v02/
matmul_60bits_vs_linbox.dat and pmbasis_app_int_intgeo_60bit_nofft.dat),
moduli above the three CRT primes (inputs reduced first) and composite moduli.
Checked against the oracle's schoolbook product, which needs no roots."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

P_NOFFT_MUL = 723501827743829161    # 60-bit non-FFT prime of the reference's matmul benchmark
P_NOFFT_PMB = 674830283164631843    # 60-bit non-FFT prime of the reference's PM-Basis benchmark
P_BIG = 4611686018427387847         # largest prime below 2^62 (above every CRT prime)
COMPOSITES = [2, 2**61, 10**18 + 9 * 7, 3 * 5 * 7 * 11 * 13 * 17 * 19 * 23 * 29 * 31 * 37 * 41 * 43 * 47,
              2**63, 2**64 - 1]
# moduli of 63 and 64 bits (the reference's tests draw primes up to 64 bits,
# t-mulmid.c:53, t-mul_geometric.c:60): every product takes the CRT path and
# the recombination reduces with a division by the normalised modulus
WORD_PRIMES = [4611686018427388039, 9223372036854775783, 9223372036854775837,
               18446744073709551557]
SMALL_NOFFT = [3, 1000003, 2**31 - 1]


@pytest.fixture(scope="module")
def pml():
    import pml_amd
    return pml_amd


_ctxs = {}


def ctx_for(pml, p):
    if p not in _ctxs:
        _ctxs[p] = pml.Context(p)
    return _ctxs[p]


CASES = [(1, 1, 1, 1, 1), (3, 4, 2, 50, 70), (8, 8, 8, 300, 200), (2, 64, 3, 129, 1000),
         (5, 1, 6, 2000, 2)]


@pytest.mark.parametrize("p", [P_NOFFT_MUL, P_NOFFT_PMB, P_BIG] + SMALL_NOFFT + COMPOSITES +
                         WORD_PRIMES)
@pytest.mark.parametrize("case", CASES)
def test_crt_product_matches_oracle(pml, p, case):
    m, k, n, la, lb = case
    A = O.rand_uniform((m, k, la), p, la + 17)
    B = O.rand_uniform((k, n, lb), p, lb + 31)
    ctx = ctx_for(pml, p)
    for method in ("auto", "plain"):
        C = ctx.mul(A, B, method=method)
        assert np.array_equal(C, O.polmat_mul(A, B, p, method="school")), (p, case, method)


def test_crt_extreme_residues(pml):
    for p in (P_NOFFT_MUL, P_BIG) + tuple(WORD_PRIMES) + (2**64 - 1,):
        la, lb = (300, 500) if p < 2**62 else (40, 50)
        A = np.full((4, 64, la), p - 1, dtype=np.uint64)
        B = np.full((64, 3, lb), p - 1, dtype=np.uint64)
        C = ctx_for(pml, p).mul(A, B)
        assert np.array_equal(C, O.polmat_mul(A, B, p, method="school"))


@pytest.mark.parametrize("p", [P_NOFFT_MUL, P_BIG, 2**61] + WORD_PRIMES)
def test_crt_mulmid_matches_oracle(pml, p):
    rng_cases = [(3, 5, 2, 100, 300, 50, 250), (4, 4, 4, 64, 65, 64, 128), (2, 3, 2, 900, 40, 0, 2000)]
    for (m, k, n, la, lb, nlo, nhi) in rng_cases:
        A = O.rand_uniform((m, k, la), p, 5 + la)
        B = O.rand_uniform((k, n, lb), p, 6 + lb)
        R = ctx_for(pml, p).mulmid(A, B, nlo, nhi)
        assert np.array_equal(R, O.polmat_mulmid(A, B, nlo, nhi, p)), (p, m, k, n, la, lb, nlo, nhi)


def test_crt_product_at_benchmark_size_evaluation_identity(pml):
    """32x32 by 32x32, length 4096, modulo the reference benchmark's non-FFT
    prime: C(r) = A(r) B(r) at random r for every entry (exact integers)."""
    p = P_NOFFT_MUL
    A = O.rand_uniform((32, 32, 4096), p, 11)
    B = O.rand_uniform((32, 32, 4096), p, 12)
    C = ctx_for(pml, p).mul(A, B)
    assert C.shape == (32, 32, 8191)
    for r in (3, 987654321987):
        Ar = [[int(v) for v in row] for row in O.polmat_eval(A, r, p)]
        Br = [[int(v) for v in row] for row in O.polmat_eval(B, r, p)]
        Cr = O.polmat_eval(C, r, p)
        for i in range(32):
            for j in range(32):
                assert sum(Ar[i][t] * Br[t][j] for t in range(32)) % p == int(Cr[i, j])


@pytest.mark.parametrize("p", [P_NOFFT_PMB, 1000003])
def test_crt_pmbasis_matches_oracle(pml, p):
    """Orders above the leaf threshold multiply through the CRT path."""
    rng = np.random.default_rng(p % 1009)
    for (m, n, order) in [(6, 3, 100), (4, 4, 70), (9, 2, 150)]:
        F = O.rand_uniform((m, n, order), p, m * 100 + order)
        sh = rng.integers(0, 4, m).astype(np.int64)
        P_ref, s_ref = O.pmbasis(F, order, sh, p)
        P_dev, s_dev = ctx_for(pml, p).pmbasis(F, order, sh)
        assert np.array_equal(O.normalise(P_dev), O.normalise(P_ref))
        assert np.array_equal(s_dev, s_ref)
        assert O.is_approximant_basis(O.normalise(P_dev), F, order, sh, p) == 0


def test_composite_modulus_rejects_bases(pml):
    ctx = ctx_for(pml, 2**61)
    assert ctx.max_log_len == -1
    with pytest.raises(pml.PmlgError) as ei:
        ctx.pmbasis(np.zeros((2, 1, 4), dtype=np.uint64), 4, np.zeros(2, dtype=np.int64))
    assert "prime" in str(ei.value)


def test_crt_length_limit_reported(pml):
    """Beyond the CRT primes' largest transform (2^24 points)."""
    ctx = ctx_for(pml, P_NOFFT_MUL)
    A = np.zeros((1, 1, (1 << 23) + 1), dtype=np.uint64)
    A[0, 0, -1] = 1
    with pytest.raises(pml.PmlgError) as ei:
        ctx.mul(A, A)
    assert ei.value.status == 2  # PMLG_ERR_UNSUPPORTED


@pytest.mark.parametrize("p", WORD_PRIMES)
def test_bases_modulo_63_and_64_bit_primes(pml, p):
    """Approximant and kernel bases for primes of 62 to 64 bits: products by
    the CRT path, M-Basis steps in plain residues with exact 128-by-64
    reductions (the Montgomery forms' lazy sums need p < 2^62)."""
    from oracle import kernel as KO
    rng = np.random.default_rng(p % 1013)
    ctx = ctx_for(pml, p)
    for (m, n, order, iterative) in [(6, 3, 90, False), (5, 2, 40, True), (9, 4, 70, False)]:
        F = O.rand_uniform((m, n, order), p, m * 7 + order)
        sh = rng.integers(0, 4, m).astype(np.int64)
        P_ref, s_ref = (O.mbasis if iterative else O.pmbasis)(F, order, sh, p)
        P_dev, s_dev = ctx.mbasis(F, order, sh) if iterative else ctx.pmbasis(F, order, sh)
        assert np.array_equal(O.normalise(P_dev), O.normalise(P_ref)), (p, m, n, order)
        assert np.array_equal(s_dev, s_ref)
        assert O.is_approximant_basis(O.normalise(P_dev), F, order, sh, p) == 0
    F = O.rand_uniform((6, 2, 9), p, 3)
    for algo, fn in (("zls", KO.kernel_zls), ("via_approx", KO.kernel_via_approx)):
        nul_r, K_r, piv_r, sh_r = fn(F, np.zeros(6, dtype=np.int64), p)
        nul, K, piv, sh = ctx.kernel_basis(F, np.zeros(6, dtype=np.int64), algo)
        assert nul == nul_r == 4
        assert np.array_equal(O.normalise(K), O.normalise(K_r)) and np.array_equal(piv, piv_r)
