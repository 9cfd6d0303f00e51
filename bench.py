"""Benchmark: 32x32 by 32x32 polynomial-matrix products, length 4096 (degree
< 4096), over Z/pZ with p = 2^60 - 3*2^24 + 1, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]

A step is one batch of B independent products C_i = A_i * B_i run through a
batch plan (one HIP-graph replay: forward transforms of every operand at 8192
points, 8192 pointwise 32x32x32 modular matrix products per product, inverse
transforms), each product's configuration picked by the learned dispatch
(the packaged network trained on recorded timings, refined by one
measurement among the configurations it predicts within 5% of its best);
inputs and outputs stay resident in HBM.  `value` is products per second of
that batched+learned form.  Beside it: the plain GPU form (one product per
call, plain per-lane pointwise kernel), the single-call MFMA form, the batched
form under the threshold rule, and (when the refinement changed the pick) the
batched form with the network's own pick.  Two more workloads are timed in
the same run: the PM-Basis of the 64x32 order-2048 matrix of the golden
fixture (its digest checked) and the 512-product uneven batch under the
learned dispatch and the threshold rule.  Prints ONE JSON line.  The CPU
baseline is the oracle's own NTT product (a C port of the same algorithm),
built for this host with -O3 -march=native and timed on one thread and on
every core this process may use, on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

P60 = 1152921504556515329
M = K = N_ = 32
LEN = 4096
LENC = 2 * LEN - 1
NPTS = 8192
PEAK_HBM_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_c2.json")
GOLDEN_FILE = os.path.join(ROOT, "tests", "golden", "golden.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--batch", type=int, default=16, help="products per step")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="time budget of the CPU baseline sample")
    ap.add_argument("--config", type=int, default=None,
                    help="run the batch in this product configuration (0..7) instead of the "
                         "learned dispatch's pick, e.g. so that profiling runs see one set of kernels")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the comparison forms")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the PM-Basis (64x32, order 2048) and 512-product batch legs")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one plain device copy of this many MiB (reference "
                         "kernel for calibrating hardware byte counters)")
    return ap.parse_args()


def time_single(ctx, dC, dA, dB, method, steps, warmup):
    """One product per call on the context's stream (the unbatched forms)."""
    for _ in range(warmup):
        ctx.mul_dev(dC.ptr, dA.ptr, dB.ptr, M, K, N_, LEN, LEN, method)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.mul_dev(dC.ptr, dA.ptr, dB.ptr, M, K, N_, LEN, LEN, method)
    ctx.sync()
    return steps / (time.perf_counter() - t0)


def time_plan(plan, steps, warmup, staged=False):
    for _ in range(warmup):
        plan.run()
    plan.sync()
    if staged:
        plan.stage_times()  # the stage counts restart with the timed region
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
    plan.sync()
    return time.perf_counter() - t0


def splitmix_uniform(n, p, seed):
    """Residues in [0, p) from the SplitMix64 stream of `seed` (raw draw t is
    mix(seed + t * 0x9e3779b97f4a7c15), t = 1, 2, ...), each draw's top
    bitlen(p - 1) bits kept when below p: the stream the golden fixtures were
    written from, so the bench runs the fixture's own inputs."""
    bits = int(p - 1).bit_length()
    out, got, t = [], 0, 0
    with np.errstate(over="ignore"):
        while got < n:
            chunk = int((n - got) * 1.1) + 64
            z = np.uint64(seed) + np.arange(t + 1, t + chunk + 1, dtype=np.uint64) * \
                np.uint64(0x9e3779b97f4a7c15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xbf58476d1ce4e5b9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94d049bb133111eb)
            z = z ^ (z >> np.uint64(31))
            x = z >> np.uint64(64 - bits)
            x = x[x < np.uint64(p)][:n - got]
            out.append(x)
            got += len(x)
            t += chunk
    return np.concatenate(out)


def bench_c3(reps):
    """PM-Basis of the 64x32 order-2048 matrix of the golden fixture (30-bit
    prime, zero shift), F and P resident in HBM; the basis is compared with
    the fixture's digest."""
    import hashlib
    import pml_amd
    g = json.load(open(GOLDEN_FILE))["c3_pmbasis_64x32_order2048_p30"]
    m, n, order, p = g["m"], g["n"], g["order"], g["p"]
    F = splitmix_uniform(m * n * order, p, g["seedF"]).reshape(m, n, order)
    cap = order + 1
    with pml_amd.Context(p) as ctx:
        dF = pml_amd.DeviceArray(ctx, F.nbytes)
        dF.upload(F)
        dP = pml_amd.DeviceArray(ctx, m * m * cap * 8)
        sh = np.zeros(m, dtype=np.int64)
        L = ctx.pmbasis_dev(dP.ptr, cap, sh, dF.ptr, m, n, order, order)  # leaf order, arena
        t0 = time.perf_counter()
        for _ in range(reps):
            sh = np.zeros(m, dtype=np.int64)
            L = ctx.pmbasis_dev(dP.ptr, cap, sh, dF.ptr, m, n, order, order)
        dt = (time.perf_counter() - t0) / reps
        P = dP.download(np.empty((m, m, cap), dtype=np.uint64))[:, :, :L]
        while P.shape[2] and not P[:, :, -1].any():
            P = P[:, :, :-1]
        ok = hashlib.sha256(np.ascontiguousarray(P).tobytes()).hexdigest() == g["sha256"] and \
            [int(x) for x in sh] == g["shift_out"]
        leaf = ctx.leaf_order(m, n)
        dF.free()
        dP.free()
    return {"seconds_per_basis": round(dt, 5), "bases_per_s": round(1.0 / dt, 2),
            "reps": reps, "leaf_order": leaf, "golden_digest_match": bool(ok),
            "workload": "pmbasis 64x32, order 2048, p=998244353, zero shift, F and P in HBM"}


def bench_c4(reps):
    """512 uneven products (m, k, n uniform in 2..64, lengths log-uniform in
    16..16384, 30-/60-bit FFT primes 50/50, seed 4), operands generated on
    the device, one batch plan replayed as a graph: learned dispatch (with the
    measured refinement) and the threshold rule."""
    import ctypes
    from pml_amd import learn, lib, _check
    from pml_amd.perf import DeviceBatch
    shapes = learn.draw_shapes(512, 4)
    model = learn.install_default()  # the learned leg's model (C2's is closed by now)
    batch = DeviceBatch(shapes, "auto")
    out = {"workload": "512 uneven products, m,k,n in 2..64, lengths 16..16384 log-uniform, "
                       "p in {998244353, 2^60-3*2^24+1}, seed 4",
           "algorithmic_bytes_per_batch": batch.bytes, "reps": reps}
    for name, meth in (("learned", 0), ("threshold", 3)):
        for i in range(len(shapes)):
            batch.descs[i].method = meth
        t0 = time.perf_counter()
        plan = batch.plan(1)
        plan_s = time.perf_counter() - t0
        _check(lib().pmlg_batch_plan_run(plan, None))  # eager pass + capture
        _check(lib().pmlg_batch_plan_run(plan, None))
        _check(lib().pmlg_batch_plan_sync(plan))
        t0 = time.perf_counter()
        for _ in range(reps):
            _check(lib().pmlg_batch_plan_run(plan, None))
        _check(lib().pmlg_batch_plan_sync(plan))
        dt = (time.perf_counter() - t0) / reps
        launches, arena = ctypes.c_int64(), ctypes.c_uint64()
        _check(lib().pmlg_batch_plan_stats(plan, ctypes.byref(launches), ctypes.byref(arena)))
        lib().pmlg_batch_plan_destroy(plan)
        out[name] = {"seconds_per_batch": round(dt, 5), "products_per_s": round(512 / dt, 1),
                     "plan_seconds": round(plan_s, 2), "launches": launches.value}
    batch.close()
    model.close()
    return out


def freivalds(A, B, C, p, seed):
    """C(r) == A(r) B(r) mod p at a random r (every entry, exact integers)."""
    from oracle import oracle as O
    r = int(np.random.default_rng(seed).integers(2, p - 1))
    Ar = [[int(v) for v in row] for row in O.polmat_eval(A, r, p)]
    Br = [[int(v) for v in row] for row in O.polmat_eval(B, r, p)]
    Cr = O.polmat_eval(C, r, p)
    for i in range(A.shape[0]):
        for j in range(B.shape[1]):
            if sum(Ar[i][t] * Br[t][j] for t in range(A.shape[1])) % p != int(Cr[i, j]):
                return False
    return True


def cpu_baseline(A, B, budget_s):
    """The oracle's NTT product (a C port of the same algorithm: radix-2
    transforms with Shoup butterflies, exact 128-bit pointwise sums), built
    here with -O3 -march=native, on whole 32x32x32 products: one thread for
    half the budget, then every core this process may use for the other
    half (at least one product each)."""
    import tempfile
    from oracle import oracle as O
    with tempfile.TemporaryDirectory() as d:
        flags = O.use_native(d) or "-O3 -march=x86-64-v3 -fopenmp (shipped build)"
        cores = len(os.sched_getaffinity(0))
        if os.environ.get("OMP_NUM_THREADS", "").isdigit():
            cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
        res = {}
        for th in (1, cores):
            O.set_threads(th)
            t0 = time.perf_counter()
            n = 0
            while True:
                O.polmat_mul(A, B, P60, method="ntt")
                n += 1
                el = time.perf_counter() - t0
                if el >= budget_s / 2 or n >= 64:
                    break
            res[th] = (n / el, n, el)
        O.set_threads(0)
    return res, cores, flags


def stage_bytes(batch):
    """Algorithmic HBM bytes of each stage of one batch (one launch per stage):
    forward reads the coefficients and writes the evaluations, the pointwise
    stage reads both evaluation sets and writes C's, the inverse reads C's
    evaluations and writes the coefficients."""
    ent_in = M * K + K * N_
    return {
        "forward": batch * (ent_in * LEN * 8 + ent_in * NPTS * 8),
        "pointwise": batch * (ent_in * NPTS * 8 + M * N_ * NPTS * 8),
        "inverse": batch * (M * N_ * NPTS * 8 + M * N_ * LENC * 8),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: sharding this path over several GPUs "
              "is not built (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    import pml_amd
    from pml_amd import learn

    Bn = args.batch
    rng = np.random.default_rng(20240601)
    As = [rng.integers(0, P60, size=(M, K, LEN), dtype=np.uint64) for _ in range(Bn)]
    Bs = [rng.integers(0, P60, size=(K, N_, LEN), dtype=np.uint64) for _ in range(Bn)]
    ctx = pml_amd.Context(P60)
    model = learn.install_default()
    # the network's own pick, and the pick after the measured refinement among
    # the configurations it predicts within the refinement margin of its best
    pml_amd.set_refine(False)
    model_choice = pml_amd.choose_config(ctx, M, K, N_, LEN, LEN, "auto")
    pml_amd.set_refine(True)
    if args.config is None:
        chosen = pml_amd.choose_config(ctx, M, K, N_, LEN, LEN, "auto")
        plan = pml_amd.BatchPlan([(ctx, As[i], Bs[i]) for i in range(Bn)], method="auto",
                                 graph=True)
    else:
        chosen = args.config
        plan = pml_amd.BatchPlan([(ctx, As[i], Bs[i]) for i in range(Bn)], config=chosen,
                                 graph=True)
    # graph instances with events between the stages, one per timed step: the
    # stage times below come from the timed replays themselves
    plan.time_stages(min(args.steps, 256))
    el = time_plan(plan, args.steps, args.warmup, staged=True)
    value = args.steps * Bn / el
    ms, timed_replays = plan.stage_times()

    forms = {"batched_learned": value}
    dA = pml_amd.DeviceArray(ctx, As[0].nbytes)
    dB = pml_amd.DeviceArray(ctx, Bs[0].nbytes)
    dC = pml_amd.DeviceArray(ctx, M * N_ * LENC * 8)
    dA.upload(As[0])
    dB.upload(Bs[0])
    if not args.no_forms:
        n2 = max(args.steps // 2, 5)
        forms["plain_single"] = time_single(ctx, dC, dA, dB, "plain", n2, 3)
        forms["mfma_single"] = time_single(ctx, dC, dA, dB, "mfma", n2, 3)
        with pml_amd.BatchPlan([(ctx, As[i], Bs[i]) for i in range(Bn)], method="threshold",
                               graph=True) as tplan:
            forms["batched_threshold"] = n2 * Bn / time_plan(tplan, n2, 2)
        if model_choice != chosen:
            with pml_amd.BatchPlan([(ctx, As[i], Bs[i]) for i in range(Bn)], config=model_choice,
                                   graph=True) as mplan:
                forms["batched_learned_unrefined"] = n2 * Bn / time_plan(mplan, n2, 2)

    if args.copy_mib:
        nb = args.copy_mib << 20
        src, dst = pml_amd.DeviceArray(ctx, nb), pml_amd.DeviceArray(ctx, nb)
        ctx.dev_copy(dst.ptr, src.ptr, nb)
        ctx.sync()
        src.free()
        dst.free()

    # dominant kernel: stage times of the timed graph replays (HIP events
    # recorded by the graph on the stream the kernels run on); every stage of
    # this batch is one launch
    names = ["forward", "pointwise", "inverse"]
    dom = max(range(3), key=lambda i: ms[i])
    algo = stage_bytes(Bn)
    achieved = algo[names[dom]] / (ms[dom] * 1e-3) / 1e9
    # kernels of the chosen configuration (transform radices, pointwise kernel)
    fe, ie, kind, pbs = {1: (5, 4, "plain", 0), 2: (5, 4, "mfma", 0), 3: (4, 4, "plain", 0),
                         4: (4, 4, "mfma", 0), 5: (5, 5, "mfma", 0), 6: (5, 4, "mfma", 1),
                         7: (5, 5, "plain", 0)}.get(chosen, (5, 4, "mfma", 0))
    # measured HBM traffic of the same kernel variant at this batch size
    variant = {"forward": "loge%d" % fe, "inverse": "loge%d" % ie,
               "pointwise": "mfma" if kind == "mfma" and pbs == 0 else "other"}
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        key = "%s_batch%d_%s" % (names[dom], Bn, variant[names[dom]])
        if key in tr:
            traffic = tr[key]["hbm_bytes_per_launch"]
    # (the C2 product is not truncated: lenC = 8191 > 3N/4)
    kernels = {"forward": "ntt_fwd_segs_reg<u64,u64,loose,L=13,LOGE=%d,TRUNC=0>" % fe,
               "pointwise": ("pw_mfma_segs<u64,8,%d>" % (4 >> pbs)) if kind == "mfma"
               else "pw_plain_segs<u64>",
               "inverse": "ntt_inv_segs_reg<u64,u64,loose,L=13,LOGE=%d>" % ie}
    roofline = {"bound": "hbm", "kernel": kernels[names[dom]], "achieved": round(achieved, 1),
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": traffic, "avg_launch_ms": round(ms[dom], 5),
                "algorithmic_bytes_per_launch": algo[names[dom]],
                "stage_ms_per_launch": {names[i]: round(ms[i], 5) for i in range(3)},
                "stage_ms_source": "HIP events inside the timed graph replays (%d replays)"
                                   % timed_replays}
    # the pointwise stage against the int8 MFMA peak (8x8 limb products per mod-MAC)
    int8_ops = Bn * NPTS * M * K * N_ * 64 * 2
    roofline["pointwise_int8_tops"] = round(int8_ops / (ms[1] * 1e-3) / 1e12, 1)
    roofline["pointwise_int8_frac"] = round(int8_ops / (ms[1] * 1e-3) / 5.0e15, 4)

    # parity: entry (0,0) of product 0 in full against the oracle, and the
    # evaluation identity C(r) = A(r) B(r) for every entry of every product
    from oracle import oracle as O
    res = plan.results()
    ok00 = bool(np.array_equal(res[0][:1, :1], O.polmat_mul(As[0][:1], Bs[0][:, :1], P60,
                                                           method="ntt")))
    okr = all(freivalds(As[i], Bs[i], res[i], P60, 77 + i) for i in range(Bn))
    plan.close()
    model.close()

    extra = {}
    if not args.no_extra:
        extra["c3_pmbasis"] = bench_c3(3)
        extra["c4_batch512"] = bench_c4(3)

    cpu = None
    if not args.no_cpu:
        res, cores, flags = cpu_baseline(As[0], Bs[0], args.cpu_seconds)
        r1, rc = res[1], res[cores]
        cpu = {"value": round(rc[0], 5), "unit": "products/s", "cores": cores, "kind": "port",
               "sample": "whole 32x32x32 len-4096 products by the oracle's C NTT path (Shoup "
                         "butterflies, exact 128-bit pointwise sums) built %s: %d in %.1f s on "
                         "%d threads, %d in %.1f s on 1 thread" % (flags, rc[1], rc[2], cores,
                                                                   r1[1], r1[2]),
               "value_1_thread": round(r1[0], 5), "host_cores_online": os.cpu_count(),
               "cores_available_to_process": len(os.sched_getaffinity(0))}

    out = {
        "metric": "32x32 deg-4096 mod-p polmat mul/s, batched+learned vs plain GPU, bit-exact",
        "value": round(value, 3),
        "unit": "products/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded uniform residues mod p)",
        "config": {"workload": "polmat_mul 32x32*32x32, length 4096 each, p=2^60-3*2^24+1, "
                               "NTT length 8192, 8192 pointwise 32x32x32 products per product",
                   "batch": Bn,
                   "dispatch": "learned+refined" if args.config is None else "forced",
                   "config_chosen": chosen,
                   "config_model": model_choice,
                   "parity_entry00": ok00, "parity_eval_identity_all_entries": okr},
        "forms_products_per_s": {k: round(v, 3) for k, v in forms.items()},
        "vs_plain_gpu": round(value / forms["plain_single"], 3) if "plain_single" in forms else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
