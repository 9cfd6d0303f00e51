"""GPU parity of the entry-wise host transfers (pmlg_polmat_mul_entries /
pmlg_polmat_mulmid_entries): matrices given as one coefficient array per
entry, as FLINT holds them, with ragged and empty entries, sizes spanning
many staging chunks, checked against the oracle."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

P60 = 1152921504556515329
P30 = 998244353
u64p = ctypes.POINTER(ctypes.c_uint64)
i64p = ctypes.POINTER(ctypes.c_int64)


@pytest.fixture(scope="module")
def pml():
    import pml_amd
    return pml_amd


def ragged(rng, r, c, maxlen, p, empty=0.1):
    """Entries of random lengths in [0, maxlen] (some empty); returns the
    per-entry arrays and the zero-padded dense form."""
    ents, lens = [], []
    dense = np.zeros((r, c, maxlen), dtype=np.uint64)
    for i in range(r):
        for j in range(c):
            l = 0 if rng.random() < empty else int(rng.integers(1, maxlen + 1))
            a = rng.integers(0, p, l, dtype=np.uint64) if l else np.zeros(1, dtype=np.uint64)
            ents.append(a)
            lens.append(l)
            dense[i, j, :l] = a[:l]
    return ents, np.array(lens, dtype=np.int64), dense


def ptrs(arrays):
    return (u64p * len(arrays))(*[a.ctypes.data_as(u64p) for a in arrays])


def run(pml, ctx, A, B, m, k, n, nlo=None, nhi=None):
    ea, la, _ = A
    eb, lb, _ = B
    maxa, maxb = int(la.max(initial=0)), int(lb.max(initial=0))
    if nlo is None:
        nlo, nhi = 0, (maxa + maxb - 1 if maxa and maxb else 0)
    w = max(nhi - nlo, 1)
    outs = [np.full(w, 0xdead, dtype=np.uint64) for _ in range(m * n)]
    lc = np.zeros(m * n, dtype=np.int64)
    st = pml.lib().pmlg_polmat_mulmid_entries(
        ctx.handle, ptrs(outs), lc.ctypes.data_as(i64p), ptrs(ea), la.ctypes.data_as(i64p),
        ptrs(eb), lb.ctypes.data_as(i64p), m, k, n, nlo, nhi, 0)
    assert st == 0, pml.lib().pmlg_last_error()
    return outs, lc, nlo, nhi


def check(outs, lc, ref, m, n):
    for e in range(m * n):
        r = ref[e // n, e % n]
        l = len(r)
        while l > 0 and r[l - 1] == 0:
            l -= 1
        assert lc[e] == l, (e, lc[e], l)
        assert np.array_equal(outs[e][:l], r[:l])


@pytest.mark.parametrize("p", [P60, P30])
@pytest.mark.parametrize("shape", [(3, 4, 5, 50, 70), (40, 40, 40, 700, 900), (1, 64, 1, 3000, 2),
                                   (17, 3, 9, 1, 1)])
def test_entries_product(pml, p, shape):
    m, k, n, ma, mb = shape
    rng = np.random.default_rng(m * 7 + ma)
    ctx = pml.Context(p)
    A = ragged(rng, m, k, ma, p)
    B = ragged(rng, k, n, mb, p)
    outs, lc, _, _ = run(pml, ctx, A, B, m, k, n)
    check(outs, lc, O.polmat_mul(A[2], B[2], p, method="ntt"), m, n)


def test_entries_mulmid(pml):
    rng = np.random.default_rng(3)
    ctx = pml.Context(P60)
    A = ragged(rng, 6, 5, 400, P60)
    B = ragged(rng, 5, 7, 300, P60)
    for (nlo, nhi) in [(100, 300), (0, 1000), (650, 800)]:
        outs, lc, _, _ = run(pml, ctx, A, B, 6, 5, 7, nlo, nhi)
        check(outs, lc, O.polmat_mulmid(A[2], B[2], nlo, nhi, P60), 6, 7)


def test_entries_all_empty(pml):
    ctx = pml.Context(P30)
    rng = np.random.default_rng(5)
    A = ragged(rng, 3, 3, 10, P30, empty=1.0)
    B = ragged(rng, 3, 2, 10, P30)
    outs, lc, _, _ = run(pml, ctx, A, B, 3, 3, 2, 0, 19)
    assert not lc.any()
