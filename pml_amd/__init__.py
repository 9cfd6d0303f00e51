"""pml_amd: polynomial-matrix multiplication over Z/pZ on MI355X.

Python front-end of libpmlgpu.so (HIP kernels for gfx950 behind the C ABI of
include/pmlgpu.h).  It mirrors the reference's operations

    nmod_poly_mat_multiply(res, A, B)           -> Context.mul(A, B)
    nmod_poly_mat_mulmid(res, A, B, nlo, nhi)   -> Context.mulmid(A, B, nlo, nhi)
    nmod_poly_mat_pmbasis(appbas, shift, F, d)  -> Context.pmbasis(F, d, shift)
    nmod_poly_mat_mbasis(appbas, shift, F, d)   -> Context.mbasis(F, d, shift)

(flint-extras/src/nmod_poly_mat_multiply.h:92,122-124;
 flint-extras/src/nmod_poly_mat_approximant.h:282-285,321-324).

Polynomial matrices are numpy uint64 arrays of shape (rows, cols, len),
coefficients in [0, p).  There is no CPU fallback: if the library is missing
or no GPU is present, the calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmlgpu.so")

METHOD_AUTO, METHOD_PLAIN, METHOD_MFMA = 0, 1, 2
_METHODS = {"auto": METHOD_AUTO, "plain": METHOD_PLAIN, "mfma": METHOD_MFMA}

STATUS = {0: "PMLG_OK", 1: "PMLG_ERR_ARG", 2: "PMLG_ERR_UNSUPPORTED", 3: "PMLG_ERR_DEVICE",
          4: "PMLG_ERR_ALLOC"}

_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)
_LIB = None


class PmlgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS.get(status, status), msg))
        self.status = status


def lib():
    """Load libpmlgpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpmlgpu.so is not built: run `python -m pml_amd.build` "
                               "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        i64, u64, c_int, vp = ctypes.c_int64, ctypes.c_uint64, ctypes.c_int, ctypes.c_void_p
        st = ctypes.c_int
        L.pmlg_last_error.restype = ctypes.c_char_p
        L.pmlg_ctx_create.restype = st
        L.pmlg_ctx_create.argtypes = [ctypes.POINTER(vp), u64, c_int]
        L.pmlg_ctx_destroy.restype = None
        L.pmlg_ctx_destroy.argtypes = [vp]
        L.pmlg_ctx_max_log_len.restype = c_int
        L.pmlg_ctx_max_log_len.argtypes = [vp]
        L.pmlg_ctx_stream.restype = vp
        L.pmlg_ctx_stream.argtypes = [vp]
        L.pmlg_ctx_sync.restype = st
        L.pmlg_ctx_sync.argtypes = [vp]
        L.pmlg_polmat_mul.restype = st
        L.pmlg_polmat_mul.argtypes = [vp, _u64p, _u64p, _u64p, i64, i64, i64, i64, i64, c_int]
        L.pmlg_polmat_mul_dev.restype = st
        L.pmlg_polmat_mul_dev.argtypes = [vp, vp, vp, vp, i64, i64, i64, i64, i64, c_int, vp]
        L.pmlg_polmat_mulmid.restype = st
        L.pmlg_polmat_mulmid.argtypes = [vp, _u64p, _u64p, _u64p, i64, i64, i64, i64, i64, i64,
                                         i64, c_int]
        L.pmlg_polmat_mulmid_dev.restype = st
        L.pmlg_polmat_mulmid_dev.argtypes = [vp, vp, vp, vp, i64, i64, i64, i64, i64, i64, i64,
                                             c_int, vp]
        L.pmlg_profile_enable.restype = None
        L.pmlg_profile_enable.argtypes = [vp, c_int]
        L.pmlg_profile_read.restype = st
        L.pmlg_profile_read.argtypes = [vp, ctypes.POINTER(ctypes.c_double), _i64p]
        L.pmlg_dev_copy.restype = st
        L.pmlg_dev_copy.argtypes = [vp, vp, vp, u64, vp]
        L.pmlg_dev_alloc.restype = st
        L.pmlg_dev_alloc.argtypes = [vp, ctypes.POINTER(vp), u64]
        L.pmlg_dev_free.restype = None
        L.pmlg_dev_free.argtypes = [vp, vp]
        L.pmlg_memcpy_h2d.restype = st
        L.pmlg_memcpy_h2d.argtypes = [vp, vp, vp, u64]
        L.pmlg_memcpy_d2h.restype = st
        L.pmlg_memcpy_d2h.argtypes = [vp, vp, vp, u64]
        if hasattr(L, "pmlg_pmbasis"):
            L.pmlg_pmbasis.restype = st
            L.pmlg_pmbasis.argtypes = [vp, _u64p, _i64p, _i64p, _u64p, i64, i64, i64, i64, i64,
                                       c_int]
            L.pmlg_mbasis.restype = st
            L.pmlg_mbasis.argtypes = [vp, _u64p, _i64p, _i64p, _u64p, i64, i64, i64, i64, i64]
        _LIB = L
    return _LIB


def _check(status):
    if status != 0:
        raise PmlgError(status, lib().pmlg_last_error().decode())


def _p(a):
    return a.ctypes.data_as(_u64p)


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a


def _method(m):
    return _METHODS[m] if isinstance(m, str) else int(m)


class DeviceArray:
    """A raw device buffer owned by a context (for inputs resident in HBM)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        ptr = ctypes.c_void_p()
        _check(lib().pmlg_dev_alloc(ctx.handle, ctypes.byref(ptr), self.nbytes))
        self.ptr = ptr.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib().pmlg_memcpy_h2d(self.ctx.handle, self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, arr):
        assert arr.nbytes <= self.nbytes and arr.flags["C_CONTIGUOUS"]
        _check(lib().pmlg_memcpy_d2h(self.ctx.handle, arr.ctypes.data, self.ptr, arr.nbytes))
        return arr

    def free(self):
        if self.ptr:
            lib().pmlg_dev_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One prime p on one device (pmlg_ctx)."""

    def __init__(self, p, device=0):
        self.p = int(p)
        h = ctypes.c_void_p()
        _check(lib().pmlg_ctx_create(ctypes.byref(h), self.p, device))
        self.handle = h.value

    def close(self):
        if getattr(self, "handle", None):
            lib().pmlg_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def max_log_len(self):
        return lib().pmlg_ctx_max_log_len(self.handle)

    @property
    def stream(self):
        return lib().pmlg_ctx_stream(self.handle)

    def sync(self):
        _check(lib().pmlg_ctx_sync(self.handle))

    # -- products ---------------------------------------------------------
    def mul(self, A, B, method="auto"):
        """C = A*B (nmod_poly_mat_multiply); C has padded length lenA+lenB-1."""
        A, B = _c(A), _c(B)
        m, k, la = A.shape
        k2, n, lb = B.shape
        if k != k2:
            raise ValueError("inner dimensions differ")
        lc = la + lb - 1 if la > 0 and lb > 0 else 0
        C = np.zeros((m, n, lc), dtype=np.uint64)
        if lc:
            _check(lib().pmlg_polmat_mul(self.handle, _p(C), _p(A), _p(B), m, k, n, la, lb,
                                         _method(method)))
        return C

    def mulmid(self, A, B, nlo, nhi, method="auto"):
        """(A*B mod x^nhi) div x^nlo (nmod_poly_mat_mulmid), padded length nhi-nlo."""
        A, B = _c(A), _c(B)
        m, k, la = A.shape
        k2, n, lb = B.shape
        if k != k2:
            raise ValueError("inner dimensions differ")
        w = max(nhi - nlo, 0)
        out = np.zeros((m, n, w), dtype=np.uint64)
        if w:
            _check(lib().pmlg_polmat_mulmid(self.handle, _p(out), _p(A), _p(B), m, k, n, la, lb,
                                            nlo, nhi, _method(method)))
        return out

    def dev_copy(self, dst, src, nbytes, stream=None):
        _check(lib().pmlg_dev_copy(self.handle, int(dst), int(src), int(nbytes), stream))

    def mul_dev(self, dC, dA, dB, m, k, n, la, lb, method="auto", stream=None):
        """Device-pointer product, enqueued without synchronisation."""
        _check(lib().pmlg_polmat_mul_dev(self.handle, int(dC), int(dA), int(dB), m, k, n, la,
                                         lb, _method(method), stream))

    # -- approximant bases -------------------------------------------------
    def _basis(self, fn, F, order, shift, *extra):
        F = _c(F)
        m, n, lf = F.shape
        sh = np.ascontiguousarray(shift, dtype=np.int64).copy()
        cap = max(order, 0) + 1
        P = np.zeros((m, m, cap), dtype=np.uint64)
        L = ctypes.c_int64(0)
        _check(fn(self.handle, _p(P), ctypes.byref(L), sh.ctypes.data_as(_i64p), _p(F), m, n,
                  lf, order, cap, *extra))
        return P[:, :, : L.value].copy(), sh

    def pmbasis(self, F, order, shift, method="auto"):
        """Shift-ordered weak Popov approximant basis (nmod_poly_mat_pmbasis)."""
        return self._basis(lib().pmlg_pmbasis, F, order, shift, _method(method))

    def mbasis(self, F, order, shift):
        """Iterative approximant basis (nmod_poly_mat_mbasis)."""
        return self._basis(lib().pmlg_mbasis, F, order, shift)

    # -- stage timing ---------------------------------------------------------
    def profile(self, on=True):
        lib().pmlg_profile_enable(self.handle, 1 if on else 0)

    def profile_read(self):
        ms = (ctypes.c_double * 3)()
        cnt = (ctypes.c_int64 * 3)()
        _check(lib().pmlg_profile_read(self.handle, ms, cnt))
        return list(ms), list(cnt)
