"""pml_amd: polynomial-matrix multiplication over Z/pZ on MI355X.

Python front-end of libpmlgpu.so (HIP kernels for gfx950 behind the C ABI of
include/pmlgpu.h).  It mirrors the reference's operations

    nmod_poly_mat_multiply(res, A, B)           -> Context.mul(A, B)
    nmod_poly_mat_mulmid(res, A, B, nlo, nhi)   -> Context.mulmid(A, B, nlo, nhi)
    nmod_poly_mat_pmbasis(appbas, shift, F, d)  -> Context.pmbasis(F, d, shift)
    nmod_poly_mat_mbasis(appbas, shift, F, d)   -> Context.mbasis(F, d, shift)
    nmod_poly_mat_kernel_zls_approx(ker, piv, shift, F)
                                                -> Context.kernel_basis(F, shift, "zls")
    nmod_poly_mat_kernel_via_approx(ker, piv, shift, F)
                                                -> Context.kernel_basis(F, shift, "via_approx")

(flint-extras/src/nmod_poly_mat_multiply.h:92,122-124;
 flint-extras/src/nmod_poly_mat_approximant.h:282-285,321-324;
 flint-extras/src/nmod_poly_mat_kernel.h:91-105).

Polynomial matrices are numpy uint64 arrays of shape (rows, cols, len),
coefficients in [0, p).  There is no CPU fallback: if the library is missing
or no GPU is present, the calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmlgpu.so")

METHOD_AUTO, METHOD_PLAIN, METHOD_MFMA, METHOD_THRESHOLD = 0, 1, 2, 3
_METHODS = {"auto": METHOD_AUTO, "plain": METHOD_PLAIN, "mfma": METHOD_MFMA,
            "threshold": METHOD_THRESHOLD}
BATCH_GRAPH = 1
KERNEL_ZLS, KERNEL_VIA_APPROX = 0, 1


class MulDesc(ctypes.Structure):
    """pmlg_mul_desc: one product of a batch (device pointers)."""
    _fields_ = [("ctx", ctypes.c_void_p), ("A", ctypes.c_void_p), ("B", ctypes.c_void_p),
                ("C", ctypes.c_void_p), ("m", ctypes.c_int64), ("k", ctypes.c_int64),
                ("n", ctypes.c_int64), ("lenA", ctypes.c_int64), ("lenB", ctypes.c_int64),
                ("method", ctypes.c_int), ("config", ctypes.c_int)]

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
        if hasattr(L, "pmlg_polmat_mul_dev_noalloc"):
            L.pmlg_polmat_mul_dev_noalloc.restype = st
            L.pmlg_polmat_mul_dev_noalloc.argtypes = [vp, vp, vp, vp, i64, i64, i64, i64, i64,
                                                      c_int, vp]
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
        L.pmlg_batch_plan_create.restype = st
        L.pmlg_batch_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(MulDesc), i64,
                                             c_int]
        L.pmlg_batch_plan_run.restype = st
        L.pmlg_batch_plan_run.argtypes = [vp, vp]
        L.pmlg_batch_plan_sync.restype = st
        L.pmlg_batch_plan_sync.argtypes = [vp]
        L.pmlg_batch_plan_profile.restype = st
        L.pmlg_batch_plan_profile.argtypes = [vp, c_int, ctypes.POINTER(ctypes.c_double)]
        L.pmlg_batch_plan_time_stages.restype = st
        L.pmlg_batch_plan_time_stages.argtypes = [vp, c_int]
        L.pmlg_batch_plan_stage_times.restype = st
        L.pmlg_batch_plan_stage_times.argtypes = [vp, ctypes.POINTER(ctypes.c_double), _i64p]
        L.pmlg_batch_plan_config.restype = st
        L.pmlg_batch_plan_config.argtypes = [vp, i64, ctypes.POINTER(c_int)]
        L.pmlg_batch_plan_stats.restype = st
        L.pmlg_batch_plan_stats.argtypes = [vp, _i64p, _u64p]
        L.pmlg_batch_plan_destroy.restype = None
        L.pmlg_batch_plan_destroy.argtypes = [vp]
        L.pmlg_polmat_mul_batch.restype = st
        L.pmlg_polmat_mul_batch.argtypes = [ctypes.POINTER(MulDesc), i64, vp]
        L.pmlg_fill_random.restype = st
        L.pmlg_fill_random.argtypes = [vp, vp, i64, u64, vp]
        L.pmlg_num_configs.restype = c_int
        L.pmlg_num_configs.argtypes = []
        L.pmlg_config_valid.restype = c_int
        L.pmlg_config_valid.argtypes = [vp, i64, i64, i64, i64, i64, c_int]
        L.pmlg_choose_config.restype = c_int
        L.pmlg_choose_config.argtypes = [vp, i64, i64, i64, i64, i64, c_int]
        L.pmlg_autotune.restype = st
        L.pmlg_autotune.argtypes = [vp, i64, i64, i64, i64, i64, c_int,
                                    ctypes.POINTER(ctypes.c_double), ctypes.POINTER(c_int)]
        if hasattr(L, "pmlg_pmbasis"):
            L.pmlg_pmbasis.restype = st
            L.pmlg_pmbasis.argtypes = [vp, _u64p, _i64p, _i64p, _u64p, i64, i64, i64, i64, i64,
                                       c_int]
            L.pmlg_mbasis.restype = st
            L.pmlg_mbasis.argtypes = [vp, _u64p, _i64p, _i64p, _u64p, i64, i64, i64, i64, i64]
            L.pmlg_pmbasis_set_leaf_order.restype = st
            L.pmlg_pmbasis_set_leaf_order.argtypes = [vp, i64]
            L.pmlg_pmbasis_leaf_order.restype = st
            L.pmlg_pmbasis_leaf_order.argtypes = [vp, i64, i64, c_int, _i64p]
        L.pmlg_dispatch_set_refine.restype = None
        L.pmlg_dispatch_set_refine.argtypes = [c_int]
        if hasattr(L, "pmlg_polmat_mul_entries"):
            pp = ctypes.POINTER(_u64p)
            L.pmlg_polmat_mul_entries.restype = st
            L.pmlg_polmat_mul_entries.argtypes = [vp, pp, _i64p, pp, _i64p, pp, _i64p, i64, i64,
                                                  i64, c_int]
            L.pmlg_polmat_mulmid_entries.restype = st
            L.pmlg_polmat_mulmid_entries.argtypes = [vp, pp, _i64p, pp, _i64p, pp, _i64p, i64,
                                                     i64, i64, i64, i64, c_int]
        L.pmlg_pmbasis_dev.restype = st
        L.pmlg_pmbasis_dev.argtypes = [vp, vp, i64, _i64p, _i64p, vp, i64, i64, i64, i64, c_int, vp]
        if hasattr(L, "pmlg_kernel_basis"):
            L.pmlg_kernel_basis.restype = st
            L.pmlg_kernel_basis.argtypes = [vp, _u64p, i64, _i64p, _i64p, _i64p, _i64p, _u64p, i64,
                                            i64, i64, c_int, c_int]
            L.pmlg_kernel_basis_dev.restype = st
            L.pmlg_kernel_basis_dev.argtypes = [vp, vp, i64, _i64p, _i64p, _i64p, _i64p, vp, i64,
                                                i64, i64, c_int, c_int, vp]
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

    def fill_random(self, dst, n, seed, stream=None):
        """n residues in [0, p) from a seeded stream, generated on the device."""
        _check(lib().pmlg_fill_random(self.handle, int(dst), int(n), int(seed), stream))

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
        P = np.empty((m, m, cap), dtype=np.uint64)  # every coefficient is written
        L = ctypes.c_int64(0)
        _check(fn(self.handle, _p(P), ctypes.byref(L), sh.ctypes.data_as(_i64p), _p(F), m, n,
                  lf, order, cap, *extra))
        return P[:, :, : L.value].copy(), sh

    def pmbasis_dev(self, dP, capP, shift, dF, m, n, lenF, order, method="auto", stream=None):
        """PM-Basis on device buffers (pmlg_pmbasis_dev): F is m x n x lenF and P
        m x m x capP (capP >= order + 1) in device memory, entry-major; `shift`
        (int64 numpy array) is updated in place.  Returns the basis length."""
        sh = np.ascontiguousarray(shift, dtype=np.int64)
        if sh is not shift:
            raise ValueError("shift must be a contiguous int64 array (updated in place)")
        L = ctypes.c_int64(0)
        _check(lib().pmlg_pmbasis_dev(self.handle, int(dP), int(capP), ctypes.byref(L),
                                      sh.ctypes.data_as(_i64p), int(dF), m, n, lenF, order,
                                      _method(method), stream))
        return L.value

    def pmbasis(self, F, order, shift, method="auto"):
        """Shift-ordered weak Popov approximant basis (nmod_poly_mat_pmbasis)."""
        return self._basis(lib().pmlg_pmbasis, F, order, shift, _method(method))

    def mbasis(self, F, order, shift):
        """Iterative approximant basis (nmod_poly_mat_mbasis)."""
        return self._basis(lib().pmlg_mbasis, F, order, shift)

    def set_leaf_order(self, order):
        """M-Basis / PM-Basis crossover (pmlg_pmbasis_set_leaf_order): 0 measures
        it per shape for calls of order >= 512, 1..256 fixes it."""
        _check(lib().pmlg_pmbasis_set_leaf_order(self.handle, int(order)))

    def leaf_order(self, m, n, method="auto"):
        """Leaf order a PM-Basis call of order >= 512 on an m x n input uses."""
        t = ctypes.c_int64(0)
        _check(lib().pmlg_pmbasis_leaf_order(self.handle, m, n, _method(method), ctypes.byref(t)))
        return t.value

    def kernel_basis(self, F, shift=None, algo="zls", method="auto", cap=None):
        """Shift-ordered weak Popov basis of the left kernel of F
        (nmod_poly_mat_kernel_zls_approx / _via_approx).  Returns (nullity, K,
        pivind, shift_out): K has nullity rows, shift_out their shifted row
        degrees.  shift=None takes the row degrees of F (zero rows as 0)."""
        F = _c(F)
        m, n, lf = F.shape
        if shift is None:
            nz = F != 0
            deg = np.where(nz.any(axis=2), lf - 1 - np.argmax(nz[..., ::-1], axis=2), -1)
            shift = np.maximum(deg.max(axis=1), 0) if n else np.zeros(m)
        sh0 = np.ascontiguousarray(shift, dtype=np.int64).copy()
        a = {"zls": KERNEL_ZLS, "via_approx": KERNEL_VIA_APPROX}[algo] if isinstance(algo, str) \
            else int(algo)
        if not cap:
            # first guess from the generic kernel degree n d / (m - n) (the
            # shifted row degrees of a kernel basis add up to at most n d),
            # with room to spare; the call reports the exact length when the
            # guess is short and is then repeated with it
            spread = int(sh0.max() - sh0.min()) if m else 0
            cap = 2 * (n * max(lf, 1) // max(m - n, 1)) + spread + 64
        cap = int(cap)
        while True:
            sh = sh0.copy()
            piv = np.zeros(m, dtype=np.int64)
            K = np.zeros((m, m, cap), dtype=np.uint64)
            L, nul = ctypes.c_int64(0), ctypes.c_int64(0)
            status = lib().pmlg_kernel_basis(self.handle, _p(K), cap, ctypes.byref(L),
                                             ctypes.byref(nul), piv.ctypes.data_as(_i64p),
                                             sh.ctypes.data_as(_i64p), _p(F), m, n, lf, a,
                                             _method(method))
            if status != 0 and L.value > cap:
                cap = L.value
                continue
            _check(status)
            k = nul.value
            return k, K[:k, :, :L.value].copy(), piv[:k].copy(), sh[:k].copy()

    # -- stage timing ---------------------------------------------------------
    def profile(self, on=True):
        lib().pmlg_profile_enable(self.handle, 1 if on else 0)

    def profile_read(self):
        ms = (ctypes.c_double * 3)()
        cnt = (ctypes.c_int64 * 3)()
        _check(lib().pmlg_profile_read(self.handle, ms, cnt))
        return list(ms), list(cnt)


# ---------------------------------------------------------------------------
# batched products
# ---------------------------------------------------------------------------
def num_configs():
    return lib().pmlg_num_configs()


def set_refine(on):
    """Turn the measured refinement of the learned choice on or off
    (pmlg_dispatch_set_refine): with it, configurations predicted within 5%
    of the model's best are timed once per shape class."""
    lib().pmlg_dispatch_set_refine(1 if on else 0)


class BatchPlan:
    """A planned batch of independent products (pmlg_batch_plan).

    `items` is a list of (context, A, B) with A, B numpy arrays of shape
    (m, k, lenA) and (k, n, lenB); inputs are uploaded once and stay resident,
    outputs are read back by results().  `method` / `config` apply to every
    product unless a per-item sequence is given.
    """

    def __init__(self, items, method="auto", config=-1, graph=True):
        self.items = []
        self._bufs = []
        descs = (MulDesc * max(len(items), 1))()
        methods = method if isinstance(method, (list, tuple)) else [method] * len(items)
        configs = config if isinstance(config, (list, tuple)) else [config] * len(items)
        for i, (ctx, A, B) in enumerate(items):
            A, B = _c(A), _c(B)
            m, k, la = A.shape
            k2, n, lb = B.shape
            if k != k2:
                raise ValueError("inner dimensions differ in item %d" % i)
            lc = la + lb - 1 if la > 0 and lb > 0 else 0
            dA = DeviceArray(ctx, max(A.nbytes, 8))
            dB = DeviceArray(ctx, max(B.nbytes, 8))
            dC = DeviceArray(ctx, max(m * n * lc * 8, 8))
            if A.nbytes:
                dA.upload(A)
            if B.nbytes:
                dB.upload(B)
            self._bufs += [dA, dB, dC]
            self.items.append((ctx, (m, k, n, la, lb, lc), dC))
            d = descs[i]
            d.ctx, d.A, d.B, d.C = ctx.handle, dA.ptr, dB.ptr, dC.ptr
            d.m, d.k, d.n, d.lenA, d.lenB = m, k, n, la, lb
            d.method = _method(methods[i])
            d.config = int(configs[i])
        h = ctypes.c_void_p()
        _check(lib().pmlg_batch_plan_create(ctypes.byref(h), descs, len(items),
                                            BATCH_GRAPH if graph else 0))
        self.handle = h.value

    def run(self, stream=None):
        _check(lib().pmlg_batch_plan_run(self.handle, stream))

    def sync(self):
        _check(lib().pmlg_batch_plan_sync(self.handle))

    def profile(self, reps=5):
        """Average ms of the forward, pointwise and inverse stages (eager runs)."""
        ms = (ctypes.c_double * 3)()
        _check(lib().pmlg_batch_plan_profile(self.handle, reps, ms))
        return list(ms)

    def time_stages(self, ring):
        """Capture `ring` graph instances with events between the stages; the
        following replays use them in turn (stage_times reads them)."""
        _check(lib().pmlg_batch_plan_time_stages(self.handle, int(ring)))

    def stage_times(self):
        """(ms per stage averaged over the timed replays since the last call,
        number of replays averaged)."""
        ms = (ctypes.c_double * 3)()
        n = ctypes.c_int64()
        _check(lib().pmlg_batch_plan_stage_times(self.handle, ms, ctypes.byref(n)))
        return list(ms), n.value

    def config(self, i):
        c = ctypes.c_int()
        _check(lib().pmlg_batch_plan_config(self.handle, i, ctypes.byref(c)))
        return c.value

    def stats(self):
        l, b = ctypes.c_int64(), ctypes.c_uint64()
        _check(lib().pmlg_batch_plan_stats(self.handle, ctypes.byref(l), ctypes.byref(b)))
        return l.value, b.value

    def result(self, i):
        ctx, (m, k, n, la, lb, lc), dC = self.items[i]
        C = np.zeros((m, n, lc), dtype=np.uint64)
        if C.nbytes:
            dC.download(C)
        return C

    def results(self):
        return [self.result(i) for i in range(len(self.items))]

    def close(self):
        if getattr(self, "handle", None):
            lib().pmlg_batch_plan_destroy(self.handle)
            self.handle = None
        for b in self._bufs:
            b.free()
        self._bufs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def mul_batch(items, method="auto", config=-1, graph=False):
    """Products of a batch through one plan run; returns the list of C_i."""
    with BatchPlan(items, method=method, config=config, graph=graph) as plan:
        plan.run(None)
        plan.sync()
        return plan.results()


def autotune(ctx, m, k, n, lenA, lenB, reps=2):
    """Times every valid configuration; returns (ms list with -1 for untimed, best)."""
    K = num_configs()
    ms = (ctypes.c_double * K)()
    best = ctypes.c_int()
    _check(lib().pmlg_autotune(ctx.handle, m, k, n, lenA, lenB, reps, ms, ctypes.byref(best)))
    return list(ms), best.value


def choose_config(ctx, m, k, n, lenA, lenB, method="auto"):
    return lib().pmlg_choose_config(ctx.handle, m, k, n, lenA, lenB, _method(method))


def config_valid(ctx, m, k, n, lenA, lenB, config):
    return bool(lib().pmlg_config_valid(ctx.handle, m, k, n, lenA, lenB, config))
