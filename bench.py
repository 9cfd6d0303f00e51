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
batched form with the network's own pick.  Prints ONE JSON line.  The CPU
baseline is the oracle's own NTT product (a single-threaded C port of the
same algorithm) timed on this host on a bounded sample.
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--batch", type=int, default=16, help="products per step")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="time budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the comparison forms")
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


def time_plan(plan, steps, warmup):
    for _ in range(warmup):
        plan.run()
    plan.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.run()
    plan.sync()
    return time.perf_counter() - t0


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
    """Oracle NTT product (single thread) on whole 32x32x32 products until the
    time budget is used (at least one)."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    n = 0
    while True:
        O.polmat_mul(A, B, P60, method="ntt")
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 8:
            break
    return n / el, n, el


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
    # the configurations it predicts within 5% of its best
    pml_amd.set_refine(False)
    model_choice = pml_amd.choose_config(ctx, M, K, N_, LEN, LEN, "auto")
    pml_amd.set_refine(True)
    chosen = pml_amd.choose_config(ctx, M, K, N_, LEN, LEN, "auto")

    plan = pml_amd.BatchPlan([(ctx, As[i], Bs[i]) for i in range(Bn)], method="auto", graph=True)
    el = time_plan(plan, args.steps, args.warmup)
    value = args.steps * Bn / el

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

    # dominant kernel: stage times of eager runs of the same plan (HIP events
    # on the plan's stream); every stage of this batch is one launch
    ms = plan.profile(reps=5)
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
    kernels = {"forward": "ntt_fwd_segs_reg<u64,u64,loose,L=13,LOGE=%d>" % fe,
               "pointwise": ("pw_mfma_segs<u64,8,%d>" % (4 >> pbs)) if kind == "mfma"
               else "pw_plain_segs<u64>",
               "inverse": "ntt_inv_segs_reg<u64,u64,loose,L=13,LOGE=%d>" % ie}
    roofline = {"bound": "hbm", "kernel": kernels[names[dom]], "achieved": round(achieved, 1),
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                "traffic": traffic, "avg_launch_ms": round(ms[dom], 5),
                "algorithmic_bytes_per_launch": algo[names[dom]],
                "stage_ms_per_launch": {names[i]: round(ms[i], 5) for i in range(3)}}
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

    cpu = None
    if not args.no_cpu:
        rate, n, el_cpu = cpu_baseline(As[0], Bs[0], args.cpu_seconds)
        cpu = {"value": round(rate, 5), "unit": "products/s", "cores": 1, "kind": "port",
               "sample": "%d whole 32x32x32 len-4096 product(s) by the oracle's C NTT path "
                         "(Shoup butterflies, 128-bit accumulation), %.1f s on one host core"
                         % (n, el_cpu)}

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
                   "batch": Bn, "dispatch": "learned+refined", "config_chosen": chosen,
                   "config_model": model_choice,
                   "parity_entry00": ok00, "parity_eval_identity_all_entries": okr},
        "forms_products_per_s": {k: round(v, 3) for k, v in forms.items()},
        "vs_plain_gpu": round(value / forms["plain_single"], 3) if "plain_single" in forms else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
