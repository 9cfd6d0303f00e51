"""The product entry point that allocates nothing (pmlg_polmat_mul_dev_noalloc),
called through the C ABI and compared bit for bit with pmlg_polmat_mul_dev
(whose products the other GPU tests compare with the oracle) and with the
oracle itself.  Two no-allocation calls in a row, the second on a smaller
input, share the device's fixed scratch array; the first leaves the array
full of its evaluations, which the second must not read.  Sizes beyond the
array, products that need the multi-prime path and transform lengths whose
root tables are not yet built are refused before any work is enqueued."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

P60 = 1152921504556515329
P30 = 998244353
P_NOFFT = 723501827743829161
SCRATCH = 256 << 20


@pytest.fixture(scope="module")
def pml():
    import pml_amd
    return pml_amd


def _upload(pml, ctx, X):
    d = pml.DeviceArray(ctx, max(X.nbytes, 8))
    d.upload(X)
    return d


def _run(pml, ctx, fn, A, B, method):
    m, k, la = A.shape
    n, lb = B.shape[1], B.shape[2]
    dA, dB = _upload(pml, ctx, A), _upload(pml, ctx, B)
    C = np.full((m, n, la + lb - 1), 7, dtype=np.uint64)
    dC = _upload(pml, ctx, C)  # stale content the call must overwrite
    st = fn(ctx.handle, dC.ptr, dA.ptr, dB.ptr, m, k, n, la, lb, method, None)
    ctx.sync()
    if st == 0:
        dC.download(C)
    for d in (dA, dB, dC):
        d.free()
    return st, C


# (p, method, first shape, second smaller shape); shapes are (m, k, n, lenA, lenB)
CASES = [
    (P60, 2, (16, 16, 16, 2000, 1500), (3, 5, 4, 300, 200)),     # MFMA, 64-bit words
    (P60, 1, (8, 8, 8, 4096, 4096), (2, 3, 2, 1, 7)),             # plain; N = 8192, then 8
    (P30, 0, (12, 7, 9, 5000, 3000), (1, 1, 1, 33, 31)),          # 32-bit words
    (P60, 0, (2, 3, 2, 9000, 7000), (2, 2, 2, 5000, 3000)),       # N = 2^14 and 2^13: merge pass
    (P30, 2, (3, 2, 2, 12000, 9000), (4, 4, 4, 64, 64)),          # N = 2^15 on 32-bit words
]


@pytest.mark.parametrize("p,method,big,small", CASES)
def test_noalloc_matches_mul_dev_two_calls(pml, p, method, big, small):
    L = pml.lib()
    rng = np.random.default_rng(p % 1000 + method)
    with pml.Context(p) as ctx:
        inputs = []
        for (m, k, n, la, lb) in (big, small):
            inputs.append((rng.integers(0, p, size=(m, k, la), dtype=np.uint64),
                           rng.integers(0, p, size=(k, n, lb), dtype=np.uint64)))
        # the allocating entry point first: its products, and the root tables
        ref = []
        for (A, B) in inputs:
            st, C = _run(pml, ctx, L.pmlg_polmat_mul_dev, A, B, method)
            assert st == 0, pml.lib().pmlg_last_error()
            ref.append(C)
        # two no-allocation calls in a row, the second smaller
        for t, (A, B) in enumerate(inputs):
            st, C = _run(pml, ctx, L.pmlg_polmat_mul_dev_noalloc, A, B, method)
            assert st == 0, L.pmlg_last_error()
            assert np.array_equal(C, ref[t]), "call %d %s" % (t, A.shape)
        # and the smaller one against the oracle
        A, B = inputs[1]
        assert np.array_equal(ref[1], O.polmat_mul(A, B, p))
        print("noalloc p=%d method=%d %s then %s: identical" % (p, method, big, small), flush=True)


def test_noalloc_refusals(pml):
    L = pml.lib()
    rng = np.random.default_rng(3)
    with pml.Context(P60) as ctx:
        # scratch: 32x32x32 at length 4096 needs 3 * 1024 * 8192 * 8 = 192 MiB (fits);
        # 48x48x48 needs 3 * 2304 * 8192 * 8 = 432 MiB (refused, PMLG_ERR_ARG)
        m = 48
        A = rng.integers(0, P60, size=(m, m, 4096), dtype=np.uint64)
        B = rng.integers(0, P60, size=(m, m, 4096), dtype=np.uint64)
        st, _ = _run(pml, ctx, L.pmlg_polmat_mul_dev, A[:1, :1], B[:1, :1], 0)  # tables of 2^13
        assert st == 0
        st, C = _run(pml, ctx, L.pmlg_polmat_mul_dev_noalloc, A, B, 0)
        assert st == 1 and b"scratch" in L.pmlg_last_error()
        assert (C == 7).all()  # nothing was written
        # root tables of 2^10 not built on this context yet
        A2 = rng.integers(0, P60, size=(2, 2, 500), dtype=np.uint64)
        st, _ = _run(pml, ctx, L.pmlg_polmat_mul_dev_noalloc, A2, A2, 0)
        assert st == 2 and b"root tables" in L.pmlg_last_error()
    with pml.Context(P_NOFFT) as ctx:
        A2 = rng.integers(0, P_NOFFT, size=(2, 2, 50), dtype=np.uint64)
        st, _ = _run(pml, ctx, L.pmlg_polmat_mul_dev_noalloc, A2, A2, 0)
        assert st == 2 and b"multi-prime" in L.pmlg_last_error()
