"""ctypes front-end of the CPU oracle (oracle/pml_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  The product package never
imports this module.

Polynomial matrices are numpy uint64 arrays of shape (rows, cols, len), entry
(i, j) coefficient d at [i, j, d], coefficients in [0, p).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build():
    """Compile the oracle shared library in place."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = _load(path)
    return _LIB


def use_native(workdir):
    """Compile the oracle with -O3 -march=native for this host into `workdir`
    and use that build from now on (the CPU baseline's timing leg).  Returns
    the flags used, or None (the shipped x86-64-v3 build stays) when the
    compiler is missing or fails."""
    global _LIB
    out = os.path.join(workdir, "liboracle_native.so")
    flags = ["-O3", "-march=native", "-fopenmp", "-fPIC", "-std=gnu11", "-shared"]
    try:
        subprocess.run(["gcc"] + flags + [os.path.join(_HERE, "pml_oracle.c"), "-o", out],
                       check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    except (OSError, subprocess.CalledProcessError):
        return None
    _LIB = _load(out)
    return " ".join(flags[:3])


def _load(path):
    L = ctypes.CDLL(path)
    u64, i64, c_int = ctypes.c_uint64, ctypes.c_int64, ctypes.c_int
    L.orc_mulmod.restype = u64
    L.orc_mulmod.argtypes = [u64, u64, u64]
    L.orc_powmod.restype = u64
    L.orc_powmod.argtypes = [u64, u64, u64]
    L.orc_invmod.restype = u64
    L.orc_invmod.argtypes = [u64, u64]
    L.orc_is_prime.restype = c_int
    L.orc_is_prime.argtypes = [u64]
    L.orc_root_of_unity_2k.restype = u64
    L.orc_root_of_unity_2k.argtypes = [u64, c_int]
    L.orc_rand_uniform.restype = None
    L.orc_rand_uniform.argtypes = [_u64p, i64, u64, u64]
    for name in ("orc_polmat_mul_school", "orc_polmat_mul"):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [_u64p, _u64p, _u64p, i64, i64, i64, i64, i64, u64]
    L.orc_polmat_mul_ntt.restype = c_int
    L.orc_polmat_mul_ntt.argtypes = [_u64p, _u64p, _u64p, i64, i64, i64, i64, i64, u64]
    L.orc_polmat_mulmid.restype = None
    L.orc_polmat_mulmid.argtypes = [_u64p, _u64p, _u64p, i64, i64, i64, i64, i64,
                                    i64, i64, u64]
    L.orc_polmat_eval.restype = None
    L.orc_polmat_eval.argtypes = [_u64p, _u64p, i64, i64, i64, u64, u64]
    L.orc_rref.restype = i64
    L.orc_rref.argtypes = [_u64p, i64, i64, u64, _i64p]
    L.orc_rank.restype = i64
    L.orc_rank.argtypes = [_u64p, i64, i64, u64]
    L.orc_left_nullspace_compact.restype = i64
    L.orc_left_nullspace_compact.argtypes = [_u64p, _i64p, _u64p, i64, i64, u64]
    for name in ("orc_mbasis", "orc_pmbasis"):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [ctypes.POINTER(_u64p), _i64p, _i64p, _u64p, i64, i64, i64, i64, u64]
    L.orc_is_approximant_basis.restype = c_int
    L.orc_is_approximant_basis.argtypes = [_u64p, i64, _u64p, i64, i64, i64, i64,
                                           _i64p, u64, u64]
    L.orc_free.restype = None
    L.orc_free.argtypes = [ctypes.c_void_p]
    L.orc_set_threads.restype = None
    L.orc_set_threads.argtypes = [c_int]
    L.orc_threads.restype = c_int
    L.orc_threads.argtypes = []
    return L


def _p(a):
    return a.ctypes.data_as(_u64p)


def _pi(a):
    return a.ctypes.data_as(_i64p)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.uint64)


def set_threads(n):
    """Host threads of the oracle's product, transform and evaluation loops
    (n <= 0: OpenMP's default, OMP_NUM_THREADS or all cores)."""
    lib().orc_set_threads(int(n))


def threads():
    return lib().orc_threads()


def mulmod(a, b, p):
    return lib().orc_mulmod(a, b, p)


def powmod(a, e, p):
    return lib().orc_powmod(a, e, p)


def is_prime(n):
    return bool(lib().orc_is_prime(n))


def root_of_unity_2k(p, k):
    return lib().orc_root_of_unity_2k(p, k)


def rand_uniform(shape, p, seed):
    out = np.empty(int(np.prod(shape)), dtype=np.uint64)
    lib().orc_rand_uniform(_p(out), out.size, p, seed)
    return out.reshape(shape)


def polmat_mul(A, B, p, method="auto"):
    A, B = _c(A), _c(B)
    m, k, la = A.shape
    k2, n, lb = B.shape
    assert k == k2
    lc = la + lb - 1 if la > 0 and lb > 0 else 0
    C = np.zeros((m, n, lc), dtype=np.uint64)
    if lc == 0:
        return C
    if method == "school":
        lib().orc_polmat_mul_school(_p(C), _p(A), _p(B), m, k, n, la, lb, p)
    elif method == "ntt":
        if lib().orc_polmat_mul_ntt(_p(C), _p(A), _p(B), m, k, n, la, lb, p) != 0:
            raise ValueError("prime does not support the NTT length")
    else:
        lib().orc_polmat_mul(_p(C), _p(A), _p(B), m, k, n, la, lb, p)
    return C


def polmat_mulmid(A, B, nlo, nhi, p):
    A, B = _c(A), _c(B)
    m, k, la = A.shape
    _, n, lb = B.shape
    lo = max(nhi - nlo, 0)
    out = np.zeros((m, n, lo), dtype=np.uint64)
    if lo:
        lib().orc_polmat_mulmid(_p(out), _p(A), _p(B), m, k, n, la, lb, nlo, nhi, p)
    return out


def polmat_eval(A, x, p):
    A = _c(A)
    r, c, ln = A.shape
    out = np.zeros((r, c), dtype=np.uint64)
    lib().orc_polmat_eval(_p(out), _p(A), r, c, ln, x, p)
    return out


def rref(M, p):
    M = _c(M).copy()
    r, c = M.shape
    piv = np.zeros(max(r, 1), dtype=np.int64)
    rank = lib().orc_rref(_p(M), r, c, p, _pi(piv))
    return M, list(piv[:rank])


def rank(M, p):
    M = _c(M)
    return lib().orc_rank(_p(M), M.shape[0], M.shape[1], p)


def left_nullspace_compact(A, p):
    A = _c(A)
    m, n = A.shape
    X = np.zeros(max(m * m, 1), dtype=np.uint64)
    perm = np.zeros(max(m, 1), dtype=np.int64)
    nullity = lib().orc_left_nullspace_compact(_p(X), _pi(perm), _p(A), m, n, p)
    rk = m - nullity
    return nullity, X[: nullity * rk].reshape(nullity, rk), list(perm[:m])


def _basis(fn, F, order, shift, p):
    F = _c(F)
    m, n, lf = F.shape
    sh = np.ascontiguousarray(shift, dtype=np.int64).copy()
    out = _u64p()
    L = ctypes.c_int64(0)
    fn(ctypes.byref(out), ctypes.byref(L), _pi(sh), _p(F), m, n, lf, order, p)
    P = np.ctypeslib.as_array(out, shape=(m * m * L.value,)).copy().reshape(m, m, L.value)
    lib().orc_free(out)
    return P, sh


def mbasis(F, order, shift, p):
    return _basis(lib().orc_mbasis, F, order, shift, p)


def pmbasis(F, order, shift, p):
    return _basis(lib().orc_pmbasis, F, order, shift, p)


def is_approximant_basis(P, F, order, shift, p, seed=1):
    P, F = _c(P), _c(F)
    m, n, lf = F.shape
    sh = np.ascontiguousarray(shift, dtype=np.int64)
    return lib().orc_is_approximant_basis(_p(P), P.shape[2], _p(F), m, n, lf, order,
                                          _pi(sh), p, seed)


def normalise(A):
    """Drop trailing all-zero coefficient slices (max normalised length)."""
    A = np.asarray(A)
    L = A.shape[-1]
    while L > 0 and not A[..., L - 1].any():
        L -= 1
    return A[..., :L]
