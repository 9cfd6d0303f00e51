"""GPU parity of the minimal kernel bases (pmlg_kernel_basis) with the oracle.

Follows the reference's kernel test (flint-extras/src/nmod_poly_mat_extra/
test/t-kernel.c:28-100, 380-700): matrix families zero / uniform / unbalanced
row and column degrees / sparse / rank-deficient, the eight shift families of
testing_collection.h, is_kernel and the returned row degrees and pivot
indices; on top of that the device basis, pivots and shifted degrees must be
bit-identical to the oracle's restatement of kernel.c (same approximant bases
underneath, so the same basis, not just an equivalent one).
"""
import ctypes
import zlib

import numpy as np
import pytest

from oracle import oracle as O
from oracle import kernel as KO


def _seed(*names):
    # stable across processes (str hashes are salted per process)
    return zlib.crc32("/".join(names).encode())


pytestmark = pytest.mark.gpu

P60 = 1152921504556515329
P30 = 998244353
P_NOFFT = 674830283164631843   # 60-bit prime without large 2-power roots (reference benchmark)
PRIMES = [P30, P60, 7340033, P_NOFFT]

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
KINDS = ["zero", "uniform", "rdeg", "cdeg", "sparse", "rkdef"]


@pytest.fixture(scope="module")
def pml():
    import pml_amd
    return pml_amd


_ctxs = {}


def ctx_for(pml, p):
    if p not in _ctxs:
        _ctxs[p] = pml.Context(p)
    return _ctxs[p]


def family(rng, kind, m, n, ln, p):
    seed = int(rng.integers(1 << 30))
    if kind == "zero":
        return np.zeros((m, n, ln), dtype=np.uint64)
    F = O.rand_uniform((m, n, ln), p, seed)
    if kind == "uniform":
        return F
    if kind == "sparse":
        return F * (rng.random((m, n, ln)) < 0.1)
    if kind == "rdeg":
        for i in range(m):
            F[i, :, int(rng.integers(1, ln + 1)):] = 0
        return F
    if kind == "cdeg":
        for j in range(n):
            F[:, j, int(rng.integers(1, ln + 1)):] = 0
        return F
    if kind == "rkdef":
        r = max(1, min(m, n) - 1)
        lg = max(1, (ln + 1) // 2)
        G = O.rand_uniform((m, r, lg), p, seed + 1)
        H = O.rand_uniform((r, n, ln - lg + 1), p, seed + 2)
        return O.polmat_mul(G, H, p)
    raise ValueError(kind)


ORACLE = {"zls": KO.kernel_zls, "via_approx": KO.kernel_via_approx}


def check(pml, p, F, shift, algo, certify=True):
    nul_r, K_r, piv_r, sh_r = ORACLE[algo](F, shift, p)
    nul, K, piv, sh = ctx_for(pml, p).kernel_basis(F, shift, algo)
    assert nul == nul_r, (algo, nul, nul_r)
    assert np.array_equal(piv, piv_r)
    assert np.array_equal(sh, sh_r)
    assert np.array_equal(O.normalise(K), O.normalise(K_r))
    if certify:
        assert KO.is_kernel(K, shift, F, p)
        if nul:
            assert np.array_equal(KO.row_degree(K, shift), sh)
            assert np.array_equal(KO.pivot_index(K, shift)[0], piv)
    return nul, K


@pytest.mark.parametrize("algo", ["zls", "via_approx"])
@pytest.mark.parametrize("kind", KINDS)
def test_kernel_matrix_families(pml, algo, kind):
    rng = np.random.default_rng(_seed(algo, kind))
    for it in range(4):
        p = PRIMES[it % len(PRIMES)]
        m, n = int(rng.integers(2, 9)), int(rng.integers(1, 7))
        F = family(rng, kind, m, n, int(rng.integers(1, 8)), p)
        check(pml, p, F, rng.integers(0, 3, m).astype(np.int64), algo)


@pytest.mark.parametrize("algo", ["zls", "via_approx"])
@pytest.mark.parametrize("shift_name", sorted(SHIFTS))
def test_kernel_shift_families(pml, algo, shift_name):
    rng = np.random.default_rng(_seed(algo, shift_name))
    for it in range(3):
        p = PRIMES[it % len(PRIMES)]
        m, n = int(rng.integers(2, 9)), int(rng.integers(1, 5))
        ln = int(rng.integers(1, 6))
        F = family(rng, ("uniform", "rkdef", "rdeg")[it], m, n, ln, p)
        check(pml, p, F, SHIFTS[shift_name](m, ln).astype(np.int64), algo)


def test_kernel_prime_two(pml):
    """Kernel bases modulo 2, both algorithms, bit-identical to the oracle."""
    rng = np.random.default_rng(22)
    for algo in ("zls", "via_approx"):
        for (m, n, ln, kind) in [(6, 2, 9, "uniform"), (8, 3, 5, "rkdef"), (5, 4, 7, "rdeg")]:
            F = family(rng, kind, m, n, ln, 2)
            shift = rng.integers(0, 3, m).astype(np.int64)
            # the oracle's is_kernel ranks by evaluation, which Z/2Z has too
            # few points for: check K F = 0, degrees and pivots directly
            nul, K = check(pml, 2, F, shift, algo, certify=False)
            if nul:
                assert not O.polmat_mul(K, F, 2).any()
                _, _, piv, sh = ctx_for(pml, 2).kernel_basis(F, shift, algo)
                assert np.array_equal(KO.row_degree(K, shift), sh)
                assert np.array_equal(KO.pivot_index(K, shift)[0], piv)


def test_kernel_row_degree_shift(pml):
    """shift = rdeg(F), the reference profile's default (p-kernel.c:75-77)."""
    rng = np.random.default_rng(3)
    for (m, n, ln) in [(6, 2, 10), (8, 3, 7), (10, 4, 5), (5, 4, 9)]:
        F = family(rng, "rdeg", m, n, ln, P60)
        shift = np.maximum(KO.degrees(F).max(axis=1), 0)
        check(pml, P60, F, shift, "zls")


def test_kernel_larger_generic(pml):
    """Generic 24 x 12 of degree 40 (products through the transforms): a
    minimal basis of 12 rows whose degrees add up to 12 * 40."""
    F = O.rand_uniform((24, 12, 41), P60, 77)
    zero = np.zeros(24, dtype=np.int64)
    for algo in ("zls", "via_approx"):
        nul, K = check(pml, P60, F, zero, algo, certify=False)
        assert nul == 12
        assert int(KO.row_degree(K, zero).sum()) == 12 * 40
        assert not O.polmat_mul(K, F, P60).any()


def test_kernel_wide_input(pml):
    """More columns than rows: the column splitting of kernel.c:252-313."""
    rng = np.random.default_rng(5)
    for (m, n) in [(5, 9), (8, 8), (3, 7)]:
        F = family(rng, "rkdef", m, n, 4, P30)
        check(pml, P30, F, np.zeros(m, dtype=np.int64), "zls")


def test_kernel_dev_entry_point(pml):
    """pmlg_kernel_basis_dev on device buffers gives the host entry's result."""
    p = P60
    ctx = ctx_for(pml, p)
    F = O.rand_uniform((7, 3, 6), p, 99)
    shift = np.zeros(7, dtype=np.int64)
    nul_h, K_h, piv_h, sh_h = ctx.kernel_basis(F, shift, "zls")
    cap = K_h.shape[2] + 3
    dF = pml.DeviceArray(ctx, F.nbytes)
    dK = pml.DeviceArray(ctx, 7 * 7 * cap * 8)
    dF.upload(F)
    sh = shift.copy()
    piv = np.zeros(7, dtype=np.int64)
    L, nul = ctypes.c_int64(0), ctypes.c_int64(0)
    i64p = ctypes.POINTER(ctypes.c_int64)
    st = pml.lib().pmlg_kernel_basis_dev(ctx.handle, dK.ptr, cap, ctypes.byref(L), ctypes.byref(nul),
                                         piv.ctypes.data_as(i64p), sh.ctypes.data_as(i64p), dF.ptr,
                                         7, 3, 6, pml.KERNEL_ZLS, 0, None)
    assert st == 0
    K = dK.download(np.zeros((7, 7, cap), dtype=np.uint64))
    assert nul.value == nul_h and L.value == K_h.shape[2]
    assert np.array_equal(K[:nul_h, :, :L.value], K_h)
    assert not K[:nul_h, :, L.value:].any()
    assert np.array_equal(piv[:nul_h], piv_h) and np.array_equal(sh[:nul_h], sh_h)
    dF.free()
    dK.free()


def test_kernel_capacity_protocol(pml):
    ctx = ctx_for(pml, P30)
    F = O.rand_uniform((6, 2, 9), P30, 5)
    nul, K, piv, sh = ctx.kernel_basis(F, np.zeros(6, dtype=np.int64), "zls", cap=1)
    nul2, K2, _, _ = ctx.kernel_basis(F, np.zeros(6, dtype=np.int64), "zls")
    assert nul == nul2 and np.array_equal(K, K2)


def test_kernel_rejects_bad_arguments(pml):
    ctx = ctx_for(pml, P30)
    with pytest.raises(pml.PmlgError):  # more than 256 rows
        ctx.kernel_basis(np.zeros((257, 1, 2), dtype=np.uint64), np.zeros(257, dtype=np.int64))
    with pytest.raises(pml.PmlgError):
        ctx.kernel_basis(np.ones((2, 130, 1), dtype=np.uint64), np.zeros(2, dtype=np.int64),
                         "via_approx")
    comp = ctx_for(pml, 2**61)
    with pytest.raises(pml.PmlgError) as ei:
        comp.kernel_basis(np.ones((3, 1, 2), dtype=np.uint64), np.zeros(3, dtype=np.int64))
    assert "prime" in str(ei.value)


@pytest.mark.parametrize("m,n,ln,p,algo,kind", [
    (100, 70, 3, P30, "zls", "uniform"),
    (90, 66, 2, P60, "via_approx", "uniform"),
    (128, 80, 2, P30, "zls", "uniform"),
    (96, 72, 3, P60, "zls", "rkdef"),
])
def test_kernel_many_columns(pml, m, n, ln, p, algo, kind):
    # more than 64 columns: the approximant bases inside hold two rows of the
    # transposed residual per lane of the composing wave
    rng = np.random.default_rng(m * 131 + n)
    F = family(rng, kind, m, n, ln, p)
    check(pml, p, F, np.zeros(m, dtype=np.int64), algo)


def test_kernel_basis_with_more_than_128_rows(pml):
    """150 and 256 rows (the approximant bases inside take up to 256 rows)."""
    for (m, n, d, p, algo) in [(150, 40, 3, P60, "zls"), (150, 40, 3, P60, "via_approx"),
                               (256, 20, 2, P30, "zls")]:
        F = O.rand_uniform((m, n, d + 1), p, m + n + d)
        nul, K = check(pml, p, F, np.zeros(m, dtype=np.int64), algo)
        assert nul == m - n
