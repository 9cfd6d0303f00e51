"""Long products: transform lengths above four times the LDS limit (2^13
points for 64-bit words, 2^14 for p < 2^30), up to 2^24 / the prime's
2-adic valuation.  The transform runs a column pass (R-point transforms
across the M columns of each entry, then the twist w^(i1 g)) before the
register kernels transform the rows, and the inverse mirrors it; see
pml_amd/csrc/ntt.hip.  The reference multiplies any length (mul.c:57-68);
each product here is compared with the oracle in full where the oracle's
NTT product runs (FFT primes), else through the evaluation identity
C(r) = A(r) B(r) at seeded points (the multi-prime path, 63/64-bit moduli)."""
import random

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

P60 = 1152921504556515329   # v2(p - 1) = 24
P30 = 998244353             # v2(p - 1) = 23
P_NOFFT = 723501827743829161
P64 = 18446744073709551557


@pytest.fixture(scope="module")
def pml():
    import pml_amd
    return pml_amd


_ctxs = {}


def ctx_for(pml, p):
    if p not in _ctxs:
        _ctxs[p] = pml.Context(p)
    return _ctxs[p]


def operands(p, m, k, n, la, lb, seed):
    return O.rand_uniform((m, k, la), p, seed), O.rand_uniform((k, n, lb), p, seed + 1)


def eval_identity(A, B, C, p, seeds=(5, 6)):
    for sd in seeds:
        r = random.Random(sd).randrange(2, p - 1)  # p may exceed 2^63
        Ar = O.polmat_eval(A, r, p)[:, :, None]
        Br = O.polmat_eval(B, r, p)[:, :, None]
        if not np.array_equal(O.polmat_mul(Ar, Br, p, "school")[:, :, 0], O.polmat_eval(C, r, p)):
            return False
    return True


FULL = [
    # p, m, k, n, lenA, lenB, method   (transform length in the comment)
    (P60, 2, 2, 2, 100000, 100000, "auto"),   # 2^18: R = 32 columns groups of 16
    (P60, 2, 2, 2, 100000, 100000, "plain"),
    (P60, 1, 3, 2, 70000, 3000, "mfma"),      # 2^17
    (P30, 2, 1, 2, 300000, 250000, "auto"),   # 2^20 on 32-bit words: R = 64
    (P60, 3, 2, 3, 40000, 30000, "mfma"),     # 2^17, several entries per operand
]


@pytest.mark.parametrize("case", FULL)
def test_long_product_matches_oracle(pml, case):
    p, m, k, n, la, lb, method = case
    A, B = operands(p, m, k, n, la, lb, la + lb)
    C = ctx_for(pml, p).mul(A, B, method=method)
    assert C.shape == (m, n, la + lb - 1)
    assert np.array_equal(C, O.polmat_mul(A, B, p, "ntt")), case
    print("long product %s identical" % (case,), flush=True)


def test_longest_transforms_match_oracle(pml):
    """The largest transform of each word size: 2^24 points for the 60-bit
    prime (R = 2^11, four columns per workgroup) and 2^23 for the 30-bit one
    (its 2-adic valuation)."""
    for p, L in ((P60, 1 << 23), (P30, 1 << 22)):
        A, B = operands(p, 1, 1, 1, L, L, 77)
        C = ctx_for(pml, p).mul(A, B)
        assert np.array_equal(C, O.polmat_mul(A, B, p, "ntt")), p
        print("length %d product identical (p=%d)" % (L, p), flush=True)


def test_long_mulmid_matches_oracle(pml):
    for p in (P60, P30):
        A, B = operands(p, 2, 2, 1, 50000, 120000, 9)
        for (nlo, nhi) in ((40000, 150000), (0, 169999), (100000, 100500)):
            R = ctx_for(pml, p).mulmid(A, B, nlo, nhi)
            assert np.array_equal(R, O.polmat_mulmid(A, B, nlo, nhi, p)), (p, nlo, nhi)


@pytest.mark.parametrize("p", [P_NOFFT, P64, 786433])
def test_long_multiprime_evaluation_identity(pml, p):
    """Moduli whose own roots do not reach the length (786433 - 1 = 3 * 2^18,
    the reference's non-FFT prime, a 64-bit prime): the three CRT primes'
    long transforms."""
    A, B = operands(p, 2, 2, 2, 60000, 70000, 3)
    C = ctx_for(pml, p).mul(A, B)
    assert C.shape == (2, 2, 129999)
    assert eval_identity(A, B, C, p)


def test_length_beyond_the_largest_transform_is_refused(pml):
    ctx = ctx_for(pml, P60)
    L = (1 << 23) + 1   # product length 2^24 + 1: transform 2^25
    A = np.zeros((1, 1, L), dtype=np.uint64)
    A[0, 0, -1] = 1
    with pytest.raises(pml.PmlgError) as ei:
        ctx.mul(A, A)
    assert ei.value.status == 2  # PMLG_ERR_UNSUPPORTED
