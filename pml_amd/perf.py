"""Timing harness for the approximant-basis and uneven-batch workloads.

    python -m pml_amd.perf pmbasis [--reps 3] [--p P]
        PM-Basis of a 64x32 matrix, order 2048, zero shift, p = 998244353 by
        default (any prime: non-FFT primes multiply through the CRT path).
    python -m pml_amd.perf mul [--p P]
        32x32 by 32x32 products of length 4096 modulo p, device-resident
        (the reference's published benchmark uses p = 723501827743829161).
    python -m pml_amd.perf batch [--pairs 512] [--method auto|threshold|plain|mfma]
        512 independent products, m, k, n uniform in 2..64, lengths
        log-uniform in 16..16384, moduli 998244353 / 2^60-3*2^24+1 (50/50),
        inputs generated on the device, one batch plan replayed as a graph.

    python -m pml_amd.perf kernel [--shapes 32x17x1024,...] [--p P] [--algo zls]
        minimal left kernel bases of random rdim x cdim matrices of degree deg,
        zero shift, p = 820293421056153991 by default (the prime of the
        reference's published kernel table, ntl-extras/benchmarks/v01/
        kernel_zls_app_int.dat); one JSON line per shape.

Each prints one JSON line.  The outputs are checked on the device where the
check is cheap (the PM-Basis residual P*F = 0 mod x^order, K*F = 0 for kernel
bases); the batch workload is checked against the oracle by the test suite.
"""
import argparse
import ctypes
import json
import math
import sys
import time

import numpy as np

from . import BatchPlan, Context, DeviceArray, MulDesc, _check, lib, learn

P30 = 998244353
P60 = 1152921504556515329


def run_mul(a):
    """One 32x32 by 32x32 product of length 4096 per call, device-resident."""
    p = a.p
    ctx = Context(p)
    M = K = N = 32
    L = 4096
    dA, dB = DeviceArray(ctx, M * K * L * 8), DeviceArray(ctx, K * N * L * 8)
    dC = DeviceArray(ctx, M * N * (2 * L - 1) * 8)
    ctx.fill_random(dA.ptr, M * K * L, 1)
    ctx.fill_random(dB.ptr, K * N * L, 2)
    ctx.mul_dev(dC.ptr, dA.ptr, dB.ptr, M, K, N, L, L)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(a.reps):
        ctx.mul_dev(dC.ptr, dA.ptr, dB.ptr, M, K, N, L, L)
    ctx.sync()
    dt = (time.perf_counter() - t0) / a.reps
    out = {"workload": "polmat_mul 32x32x32 len 4096 p=%d" % p,
           "path": "transforms mod p" if ctx.max_log_len >= 13 else "3-prime CRT",
           "seconds_per_product": round(dt, 6), "products_per_s": round(1.0 / dt, 1), "reps": a.reps}
    print(json.dumps(out))
    return 0


def run_pmbasis(a):
    m, n, order = (int(x) for x in a.shape.split("x"))
    p = a.p
    ctx = Context(p)
    ctx.set_leaf_order(a.leaf_order)
    rng = np.random.default_rng(401)
    F = rng.integers(0, p, size=(m, n, order), dtype=np.uint64)
    shift = np.zeros(m, dtype=np.int64)
    P, sh = ctx.pmbasis(F, order, shift)  # warm-up (tables, arena)
    t0 = time.perf_counter()
    for _ in range(a.reps):
        P, sh = ctx.pmbasis(F, order, shift)
    dt = (time.perf_counter() - t0) / a.reps
    # the same basis with F and P resident in device memory
    cap = order + 1
    dF = DeviceArray(ctx, F.nbytes)
    dF.upload(F)
    dP = DeviceArray(ctx, m * m * cap * 8)
    sh_d = shift.copy()
    ctx.pmbasis_dev(dP.ptr, cap, sh_d, dF.ptr, m, n, order, order)  # returns after a stream sync
    t0 = time.perf_counter()
    for _ in range(a.reps):
        sh_d = shift.copy()
        L = ctx.pmbasis_dev(dP.ptr, cap, sh_d, dF.ptr, m, n, order, order)
    dt_dev = (time.perf_counter() - t0) / a.reps
    D = dP.download(np.empty((m, m, cap), dtype=np.uint64))[:, :, :L]  # computed length
    Lh = P.shape[2]
    same = Lh <= L and np.array_equal(sh_d, sh) and np.array_equal(D[:, :, :Lh], P) and \
        not D[:, :, Lh:].any()
    # residual P*F mod x^order on the device
    R = ctx.mulmid(P, F, 0, order)
    out = {"workload": "pmbasis %dx%d order %d p=%d shift 0" % (m, n, order, p), "seconds_per_basis": round(dt, 5),
           "bases_per_s": round(1.0 / dt, 3), "seconds_per_basis_device_resident": round(dt_dev, 5),
           "device_resident_identical": bool(same), "basis_length": int(P.shape[2]),
           "shift_out_sum": int(sh.sum()), "residual_zero": bool(not R.any()), "reps": a.reps,
           "leaf_order": ctx.leaf_order(m, n)}
    print(json.dumps(out))
    return 0


class DeviceBatch:
    """The uneven-batch workload resident on the device: shapes from
    learn.draw_shapes, operands generated on the device (pmlg_fill_random,
    seeds seed0 + 2i and seed0 + 2i + 1), one descriptor per product.
    plan(flags) returns a batch-plan handle over them (the caller destroys it)."""

    def __init__(self, shapes, method="auto", seed0=1000):
        self.shapes = shapes
        self.ctxs = {}
        for (p, *_r) in shapes:
            if p not in self.ctxs:
                self.ctxs[p] = Context(p)
        meth = {"auto": 0, "plain": 1, "mfma": 2, "threshold": 3}[method]
        self.descs = (MulDesc * max(len(shapes), 1))()
        self.bufs = []
        self.bytes = 0
        self.macs = 0
        for i, (p, m, k, n, la, lb) in enumerate(shapes):
            ctx = self.ctxs[p]
            lc = la + lb - 1
            dA, dB, dC = DeviceArray(ctx, m * k * la * 8), DeviceArray(ctx, k * n * lb * 8), \
                DeviceArray(ctx, m * n * lc * 8)
            ctx.fill_random(dA.ptr, m * k * la, seed0 + 2 * i)
            ctx.fill_random(dB.ptr, k * n * lb, seed0 + 1 + 2 * i)
            self.bufs.append((dA, dB, dC))
            d = self.descs[i]
            d.ctx, d.A, d.B, d.C = ctx.handle, dA.ptr, dB.ptr, dC.ptr
            d.m, d.k, d.n, d.lenA, d.lenB, d.method, d.config = m, k, n, la, lb, meth, -1
            self.bytes += 8 * (m * k * la + k * n * lb + m * n * lc)
            self.macs += (1 << math.ceil(math.log2(lc))) * m * k * n
        for c in self.ctxs.values():
            c.sync()

    def plan(self, flags=1):
        h = ctypes.c_void_p()
        _check(lib().pmlg_batch_plan_create(ctypes.byref(h), self.descs, len(self.shapes), flags))
        return h.value

    def operands(self, i):
        """(A, B, C) of product i, read back to the host."""
        p, m, k, n, la, lb = self.shapes[i]
        dA, dB, dC = self.bufs[i]
        A = dA.download(np.empty((m, k, la), dtype=np.uint64))
        B = dB.download(np.empty((k, n, lb), dtype=np.uint64))
        C = dC.download(np.empty((m, n, la + lb - 1), dtype=np.uint64))
        return A, B, C

    def scramble_outputs(self, seed):
        """Overwrite every output with other residues (on the device)."""
        for i, (p, m, k, n, la, lb) in enumerate(self.shapes):
            self.ctxs[p].fill_random(self.bufs[i][2].ptr, m * n * (la + lb - 1), seed + i)
        for c in self.ctxs.values():
            c.sync()

    def close(self):
        for b in self.bufs:
            for x in b:
                x.free()
        self.bufs = []
        for c in self.ctxs.values():
            c.close()
        self.ctxs = {}


def run_batch(a):
    shapes = learn.draw_shapes(a.pairs, a.seed)
    model = learn.install_default() if a.method == "auto" else None
    batch = DeviceBatch(shapes, a.method)
    t0 = time.perf_counter()
    plan = batch.plan(1)
    plan_s = time.perf_counter() - t0
    _check(lib().pmlg_batch_plan_run(plan, None))  # eager pass + capture
    _check(lib().pmlg_batch_plan_sync(plan))
    t0 = time.perf_counter()
    for _ in range(a.reps):
        _check(lib().pmlg_batch_plan_run(plan, None))
    _check(lib().pmlg_batch_plan_sync(plan))
    dt = (time.perf_counter() - t0) / a.reps
    launches, arena = ctypes.c_int64(), ctypes.c_uint64()
    _check(lib().pmlg_batch_plan_stats(plan, ctypes.byref(launches), ctypes.byref(arena)))
    cfg = np.zeros(8, dtype=np.int64)
    for i in range(len(shapes)):
        c = ctypes.c_int()
        _check(lib().pmlg_batch_plan_config(plan, i, ctypes.byref(c)))
        if c.value >= 0:
            cfg[c.value] += 1
    lib().pmlg_batch_plan_destroy(plan)
    out = {"workload": "batch of %d uneven products (m, k, n uniform in 2..64, lengths "
                       "log-uniform in 16..16384, 30-/60-bit FFT primes 50/50, seed %d)"
                       % (a.pairs, a.seed),
           "method": a.method, "seconds_per_batch": round(dt, 5),
           "products_per_s": round(a.pairs / dt, 1),
           "algorithmic_GB_per_s": round(batch.bytes / dt / 1e9, 1),
           "pointwise_modmacs_per_s": float("%.4g" % (batch.macs / dt)),
           "launches_per_batch": launches.value, "arena_GB": round(arena.value / 1e9, 2),
           "plan_seconds": round(plan_s, 3), "configs": cfg.tolist(), "reps": a.reps,
           "parity": "tests/test_gpu_batch_c4.py (the same workload against the oracle)"}
    print(json.dumps(out))
    batch.close()
    if model:
        model.close()
    return 0


P_KER = 820293421056153991


def run_kernel(a):
    ctx = Context(a.p)
    for spec in a.shapes.split(","):
        m, n, d = (int(v) for v in spec.split("x"))
        rng = np.random.default_rng(m * 10007 + n * 101 + d)
        F = rng.integers(0, a.p, size=(m, n, d + 1), dtype=np.uint64)
        shift = np.zeros(m, dtype=np.int64)
        nul, K, piv, sh = ctx.kernel_basis(F, shift, a.algo)  # warm-up
        t0 = time.perf_counter()
        for _ in range(a.reps):
            nul, K, piv, sh = ctx.kernel_basis(F, shift, a.algo)
        dt = (time.perf_counter() - t0) / a.reps
        # K * F = 0 on the device
        zero = not ctx.mul(K, F).any() if nul else True
        out = {"workload": "kernel basis %dx%d deg %d p=%d shift 0 (%s)" % (m, n, d, a.p, a.algo),
               "path": "transforms mod p" if ctx.max_log_len >= 12 else "3-prime CRT",
               "seconds": round(dt, 5), "nullity": int(nul), "basis_length": int(K.shape[2]),
               "row_degree_sum": int(sh.sum()), "kernel_product_zero": bool(zero), "reps": a.reps}
        print(json.dumps(out), flush=True)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pml_amd.perf")
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("pmbasis")
    p.add_argument("--reps", type=int, default=3)
    p.add_argument("--p", type=int, default=P30, help="modulus (default 998244353)")
    p.add_argument("--leaf-order", type=int, default=0,
                   help="M-Basis/PM-Basis crossover: 0 measures it (default), 1..256 fixes it")
    p.add_argument("--shape", default="64x32x2048", help="m x n x order of F (default 64x32x2048)")
    mu = sub.add_parser("mul")
    mu.add_argument("--reps", type=int, default=20)
    mu.add_argument("--p", type=int, default=P60, help="modulus (default 2^60-3*2^24+1)")
    b = sub.add_parser("batch")
    b.add_argument("--pairs", type=int, default=512)
    b.add_argument("--seed", type=int, default=4)
    b.add_argument("--reps", type=int, default=3)
    b.add_argument("--method", default="auto", choices=["auto", "threshold", "plain", "mfma"])
    kr = sub.add_parser("kernel")
    kr.add_argument("--shapes", default="16x5x1024,32x17x1024,64x33x1024,64x49x4096")
    kr.add_argument("--p", type=int, default=P_KER)
    kr.add_argument("--algo", default="zls", choices=["zls", "via_approx"])
    kr.add_argument("--reps", type=int, default=2)
    a = ap.parse_args(argv)
    if a.cmd == "kernel":
        return run_kernel(a)
    if a.cmd == "pmbasis":
        return run_pmbasis(a)
    if a.cmd == "mul":
        return run_mul(a)
    return run_batch(a)


if __name__ == "__main__":
    sys.exit(main())
