"""Learned dispatch for the batched products: data, training, evaluation.

    python -m pml_amd.learn collect --shapes 22000 --out timings.npz
    python -m pml_amd.learn train --data timings.npz --out dispatch_mlp.npz
    python -m pml_amd.learn bands --data timings.npz --out pml_amd/csrc/bands.inc

`collect` draws shapes of the uneven-batch workload (m, k, n uniform in
2..64, each length log-uniform in 16..16384, the modulus a 30-bit or a 60-bit
FFT prime with equal probability) and records, for each, the device time of
every valid configuration (pmlg_autotune: one warm-up and `reps` timed runs,
each product timed among up to 16 identical ones, as products share launches
in a real plan).
`train` fits the fp32 network on the device (pmlg_mlp_train: the library's own
forward, backward and Adam kernels) on log(ms) of the first shapes and
evaluates on the held-out rest: total recorded time of the configurations the
network picks vs those of the threshold rule, and vs the per-shape best.
`bands` fits the banded threshold rule (the fallback, and the baseline the
network is compared with) on the
same training shapes and writes its table for the library.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import Context, PmlgError, _check, lib, num_configs, config_valid, choose_config

P60 = 1152921504556515329
P30 = 998244353
IN, OUT, NPARAMS = 5, 8, 9224
_fp = ctypes.POINTER(ctypes.c_float)


def _setup():
    L = lib()
    if getattr(L, "_mlp_ready", False):
        return L
    vp, i64, u64, st = ctypes.c_void_p, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int
    L.pmlg_mlp_create.restype = st
    L.pmlg_mlp_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int, u64]
    L.pmlg_mlp_destroy.restype = None
    L.pmlg_mlp_destroy.argtypes = [vp]
    L.pmlg_mlp_get_params.restype = st
    L.pmlg_mlp_get_params.argtypes = [vp, _fp]
    L.pmlg_mlp_set_params.restype = st
    L.pmlg_mlp_set_params.argtypes = [vp, _fp]
    L.pmlg_mlp_forward.restype = st
    L.pmlg_mlp_forward.argtypes = [vp, _fp, i64, _fp]
    L.pmlg_mlp_loss_grad.restype = st
    L.pmlg_mlp_loss_grad.argtypes = [vp, _fp, _fp, _fp, i64, _fp, _fp]
    L.pmlg_mlp_train.restype = st
    L.pmlg_mlp_train.argtypes = [vp, _fp, _fp, _fp, i64, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_float, u64, _fp]
    L.pmlg_mlp_features.restype = None
    L.pmlg_mlp_features.argtypes = [i64, i64, i64, i64, i64, u64, _fp]
    L.pmlg_dispatch_set_model.restype = st
    L.pmlg_dispatch_set_model.argtypes = [vp]
    L._mlp_ready = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(_fp)


def features(m, k, n, lenA, lenB, p):
    x = np.zeros(IN, dtype=np.float32)
    _setup().pmlg_mlp_features(m, k, n, lenA, lenB, p, _ptr(x))
    return x


class Mlp:
    """The device network (pmlg_mlp)."""

    def __init__(self, seed=1, device=0):
        L = _setup()
        h = ctypes.c_void_p()
        _check(L.pmlg_mlp_create(ctypes.byref(h), device, seed))
        self.handle = h.value

    def close(self):
        if getattr(self, "handle", None):
            _setup().pmlg_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        out = np.zeros(NPARAMS, dtype=np.float32)
        _check(_setup().pmlg_mlp_get_params(self.handle, _ptr(out)))
        return out

    @params.setter
    def params(self, theta):
        theta = _f32(theta)
        assert theta.size == NPARAMS
        _check(_setup().pmlg_mlp_set_params(self.handle, _ptr(theta)))

    def forward(self, X):
        X = _f32(X)
        out = np.zeros((X.shape[0], OUT), dtype=np.float32)
        _check(_setup().pmlg_mlp_forward(self.handle, _ptr(X), X.shape[0], _ptr(out)))
        return out

    def loss_grad(self, X, Y, M):
        X, Y, M = _f32(X), _f32(Y), _f32(M)
        loss = ctypes.c_float()
        g = np.zeros(NPARAMS, dtype=np.float32)
        _check(_setup().pmlg_mlp_loss_grad(self.handle, _ptr(X), _ptr(Y), _ptr(M), X.shape[0],
                                           ctypes.byref(loss), _ptr(g)))
        return loss.value, g

    def train(self, X, Y, M, epochs=200, batch=256, lr=1e-3, seed=7):
        X, Y, M = _f32(X), _f32(Y), _f32(M)
        loss = ctypes.c_float()
        _check(_setup().pmlg_mlp_train(self.handle, _ptr(X), _ptr(Y), _ptr(M), X.shape[0], epochs,
                                       batch, lr, seed, ctypes.byref(loss)))
        return loss.value

    def install(self):
        """Make this network the dispatcher's model for PMLG_METHOD_AUTO."""
        _check(_setup().pmlg_dispatch_set_model(self.handle))

    @staticmethod
    def uninstall():
        _check(_setup().pmlg_dispatch_set_model(None))

    def save(self, path):
        np.savez(path, params=self.params)

    @classmethod
    def load(cls, path, device=0):
        m = cls(device=device)
        with np.load(path, allow_pickle=False) as z:
            m.params = z["params"]
        return m


# ---------------------------------------------------------------------------
# data
# ---------------------------------------------------------------------------
def draw_shapes(count, seed, max_dim=64, log_len=(4.0, 14.0)):
    """Shapes of the uneven-batch workload: (p, m, k, n, lenA, lenB)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        p = P30 if rng.integers(0, 2) == 0 else P60
        m, k, n = (int(x) for x in rng.integers(2, max_dim + 1, size=3))
        la = int(round(2 ** rng.uniform(*log_len)))
        lb = int(round(2 ** rng.uniform(*log_len)))
        out.append((p, m, k, n, la, lb))
    return out


def collect(shapes, reps=1, progress=None):
    """Per-shape device times of every configuration (-1: not valid / not timed)."""
    ctxs = {}
    T = np.full((len(shapes), OUT), -1.0)
    import pml_amd
    t0 = time.time()
    for i, (p, m, k, n, la, lb) in enumerate(shapes):
        if p not in ctxs:
            ctxs[p] = Context(p)
        ms, _ = pml_amd.autotune(ctxs[p], m, k, n, la, lb, reps=reps)
        T[i, :] = ms[:OUT]
        if progress and (i + 1) % progress == 0:
            print("collect: %d/%d shapes, %.0f s" % (i + 1, len(shapes), time.time() - t0),
                  flush=True)
    for c in ctxs.values():
        c.close()
    return T


def dataset(shapes, T):
    X = np.stack([features(m, k, n, la, lb, p) for (p, m, k, n, la, lb) in shapes])
    M = (T > 0).astype(np.float32)
    Y = np.where(T > 0, np.log(np.maximum(T, 1e-6)), 0.0).astype(np.float32)
    return X, Y, M


# ---------------------------------------------------------------------------
# banded threshold rule (the fallback, and the baseline of the learned choice)
# ---------------------------------------------------------------------------
# Bands in the style of the reference's dispatch (mul.c:57-96, bands over the
# dimension and the length): word size (p < 2^30 or not), the shorter
# operand's length, the largest dimension and the product length.  Each cell
# takes the configuration of least total recorded time over the training
# shapes that fall in it; empty cells and cells whose best configuration
# cannot run a shape fall back to the fixed rule (direct convolution for
# operands of at most 8 coefficients, else MFMA when allowed, else plain).
BAND_SHORT = (8, 32, 128, 512, 2048)  # upper bounds of min(lenA, lenB) bands 0..4, then 5
BAND_DIM = (8, 16, 32)                # upper bounds of max(m, k, n) bands 0..2, then 3
NBANDS = 2 * 6 * 4 * 4


def band_index(p, m, k, n, la, lb):
    """Cell of a shape (the same function as band_index in batch.hip)."""
    w = 0 if p < (1 << 30) else 1
    short = min(la, lb)
    sb = next((i for i, b in enumerate(BAND_SHORT) if short <= b), len(BAND_SHORT))
    dim = max(m, k, n)
    db = next((i for i, b in enumerate(BAND_DIM) if dim <= b), len(BAND_DIM))
    lc = la + lb - 1
    lg = max(0, (lc - 1).bit_length())  # ceil(log2(lc))
    lb_ = min(3, max(0, (lg - 9) // 2))
    return ((w * 6 + sb) * 4 + db) * 4 + lb_


def fit_bands(shapes, T):
    """Per cell, the configuration of least total recorded time over the
    shapes of the cell for which every candidate ran (-1: none)."""
    cells = {}
    for i, (p, m, k, n, la, lb) in enumerate(shapes):
        cells.setdefault(band_index(p, m, k, n, la, lb), []).append(i)
    table = np.full(NBANDS, -1, dtype=np.int64)
    for b, idx in cells.items():
        best = None
        for c in range(OUT):
            ts = T[idx, c]
            if (ts <= 0).any():
                continue
            if best is None or ts.sum() < best[0]:
                best = (ts.sum(), c)
        if best is not None:
            table[b] = best[1]
    return table


def write_bands(table, path, ntrain, source):
    rows = ", ".join(str(int(x)) for x in table)
    with open(path, "w") as f:
        f.write("// Generated by `python -m pml_amd.learn bands` from the recorded timings\n"
                "// (%s, the first %d shapes): the configuration of each cell of the\n"
                "// banded threshold rule (-1: the fixed rule decides).  Cell index:\n"
                "// ((word * 6 + short_band) * 4 + dim_band) * 4 + length_band.\n"
                % (source, ntrain))
        f.write("static const signed char kBandConfig[%d] = {%s};\n" % (len(table), rows))


def threshold_choice(ctx_by_p, shape):
    p, m, k, n, la, lb = shape
    return choose_config(ctx_by_p[p], m, k, n, la, lb, "threshold")


def evaluate(mlp, shapes, T, ctx_by_p):
    """Total recorded time under learned, threshold and oracle-best choices."""
    X, _, M = dataset(shapes, T)
    pred = mlp.forward(X)
    tot = {"learned": 0.0, "threshold": 0.0, "best": 0.0}
    picks = np.zeros(OUT, dtype=np.int64)
    skipped = 0
    for i, sh in enumerate(shapes):
        valid = [c for c in range(OUT) if T[i, c] > 0]
        th = threshold_choice(ctx_by_p, sh)
        if not valid or T[i, th] <= 0:
            skipped += 1
            continue
        lc = min(valid, key=lambda c: pred[i, c])
        picks[lc] += 1
        tot["learned"] += T[i, lc]
        tot["threshold"] += T[i, th]
        tot["best"] += min(T[i, c] for c in valid)
    tot["gain_vs_threshold"] = tot["threshold"] / tot["learned"] if tot["learned"] else 0.0
    tot["picks"] = picks.tolist()
    tot["shapes"] = len(shapes) - skipped
    return tot


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m pml_amd.learn")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--shapes", type=int, default=22000)
    c.add_argument("--seed", type=int, default=2024)
    c.add_argument("--reps", type=int, default=1)
    c.add_argument("--range", default=None,
                   help="start:stop, a slice of the drawn shapes (collections split over runs)")
    c.add_argument("--out", required=True)
    bd = sub.add_parser("bands")
    bd.add_argument("--data", default=None)
    bd.add_argument("--train", type=int, default=20000)
    bd.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", required=True)
    t.add_argument("--train", type=int, default=20000)
    t.add_argument("--epochs", type=int, default=300)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--out", required=True)
    t.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    if a.cmd == "collect":
        shapes = draw_shapes(a.shapes, a.seed)
        if a.range:
            lo, hi = (int(v) for v in a.range.split(":"))
            shapes = shapes[lo:hi]
        t0 = time.time()
        T = collect(shapes, reps=a.reps, progress=500)
        np.savez(a.out, shapes=np.array(shapes, dtype=np.uint64), times=T)
        print("collected %d shapes in %.0f s -> %s" % (len(shapes), time.time() - t0, a.out))
        return 0
    if a.cmd == "bands":
        shapes, T = load_timings(a.data or DEFAULT_TIMINGS)
        ntr = min(a.train, len(shapes))
        table = fit_bands(shapes[:ntr], T[:ntr])
        write_bands(table, a.out, ntr, os.path.basename(a.data or DEFAULT_TIMINGS))
        print("wrote %s: %d of %d cells fitted" % (a.out, int((table >= 0).sum()), len(table)))
        return 0
    with np.load(a.data, allow_pickle=False) as z:
        shapes = [tuple(int(v) for v in row) for row in z["shapes"]]
        T = z["times"]
    ntr = min(a.train, len(shapes) - 1)
    Xtr, Ytr, Mtr = dataset(shapes[:ntr], T[:ntr])
    mlp = Mlp(seed=11)
    t0 = time.time()
    loss = mlp.train(Xtr, Ytr, Mtr, epochs=a.epochs, lr=a.lr)
    train_s = time.time() - t0
    ctx_by_p = {P30: Context(P30), P60: Context(P60)}
    ev = evaluate(mlp, shapes[ntr:], T[ntr:], ctx_by_p)
    ev_train = evaluate(mlp, shapes[:ntr], T[:ntr], ctx_by_p)
    mlp.save(a.out)
    rep = {"train_shapes": ntr, "test_shapes": ev["shapes"], "epochs": a.epochs,
           "train_seconds": round(train_s, 2), "final_train_mse_log_ms": loss,
           "test": ev, "train": ev_train}
    print(json.dumps(rep, indent=1))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    return 0



DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_MODEL = os.path.join(DATA_DIR, "dispatch_mlp.npz")
DEFAULT_TIMINGS = os.path.join(DATA_DIR, "dispatch_timings.npz")


def load_timings(path=DEFAULT_TIMINGS):
    """Recorded (shapes, times) of the uneven-batch workload on one MI355X."""
    with np.load(path, allow_pickle=False) as z:
        shapes = [tuple(int(v) for v in row) for row in z["shapes"]]
        T = z["times"].astype(np.float64)
    return shapes, T


def install_default(device=0):
    """Load the packaged network (trained on the recorded timings) and make it
    the dispatcher's model; returns it (keep it alive while it is installed)."""
    m = Mlp.load(DEFAULT_MODEL, device=device)
    m.install()
    return m


if __name__ == "__main__":
    sys.exit(main())
