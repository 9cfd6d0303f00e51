"""Device MLP of the learned dispatch: forward pass and gradients against the
fp64 restatement (oracle/mlp.py), training on the device, and the dispatch
contract (the network only chooses among exact algorithms: products stay
bit-exact whatever it picks)."""
import numpy as np
import pytest

from oracle import mlp as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu

P60 = 1152921504556515329
P30 = 998244353


@pytest.fixture(scope="module")
def learn():
    from pml_amd import learn
    return learn


def _data(rng, rows):
    X = rng.uniform(0, 1, size=(rows, 5)).astype(np.float32)
    Y = rng.normal(size=(rows, 8)).astype(np.float32)
    M = (rng.uniform(size=(rows, 8)) < 0.7).astype(np.float32)
    return X, Y, M


def test_forward_matches_fp64(learn):
    rng = np.random.default_rng(1)
    mlp = learn.Mlp(seed=3)
    theta = mlp.params.astype(np.float64)
    X, _, _ = _data(rng, 300)
    got = mlp.forward(X)
    ref, _ = R.forward(theta, X.astype(np.float64))
    # fp32 accumulation over at most 64 terms per unit
    assert np.allclose(got, ref, rtol=1e-4, atol=1e-4 * np.abs(ref).max())
    mlp.close()


@pytest.mark.parametrize("rows", [1, 37, 256])
def test_gradients_match_fp64_oracle(learn, rows):
    rng = np.random.default_rng(rows)
    mlp = learn.Mlp(seed=5)
    theta = rng.normal(scale=0.25, size=R.n_params())
    mlp.params = theta
    theta32 = mlp.params.astype(np.float64)
    X, Y, M = _data(rng, rows)
    M[0, 0] = 1.0  # at least one valid entry
    loss, g = mlp.loss_grad(X, Y, M)
    L_ref, g_ref = R.loss_grad(theta32, X.astype(np.float64), Y.astype(np.float64),
                               M.astype(np.float64))
    assert abs(loss - L_ref) <= 1e-4 * max(1.0, abs(L_ref))
    err = np.linalg.norm(g - g_ref) / max(np.linalg.norm(g_ref), 1e-30)
    assert err < 1e-4, err
    # elementwise on the large components
    big = np.abs(g_ref) > 1e-3 * np.abs(g_ref).max()
    assert np.allclose(g[big], g_ref[big], rtol=2e-3, atol=1e-6 * np.abs(g_ref).max())
    mlp.close()


def test_training_fits_a_teacher(learn):
    """Targets from a fixed random network: device training (Adam, shuffled
    batches) must drive the masked MSE far below its starting value."""
    rng = np.random.default_rng(9)
    teacher = rng.normal(scale=0.3, size=R.n_params())
    X = rng.uniform(0, 1, size=(4096, 5))
    Y, _ = R.forward(teacher, X)
    M = (rng.uniform(size=Y.shape) < 0.8).astype(np.float32)
    mlp = learn.Mlp(seed=2)
    start, _ = mlp.loss_grad(X, Y, M)
    final = mlp.train(X, Y, M, epochs=60, batch=256, lr=2e-3)
    assert final < 0.05 * start, (start, final)
    # the reported loss is the loss of the parameters left on the device
    again, _ = mlp.loss_grad(X, Y, M)
    assert abs(again - final) <= 1e-3 * max(final, 1e-6) + 1e-6
    mlp.close()


def test_learned_dispatch_picks_argmin_and_stays_exact(learn):
    import pml_amd
    rng = np.random.default_rng(12)
    mlp = learn.Mlp(seed=21)
    mlp.params = rng.normal(scale=0.5, size=R.n_params())
    ctxs = {P60: pml_amd.Context(P60), P30: pml_amd.Context(P30)}
    try:
        mlp.install()
        pml_amd.set_refine(False)   # the model's own argmin
        shapes = [(P60, 5, 7, 3, 100, 90), (P30, 9, 2, 4, 300, 17), (P60, 2, 40, 2, 33, 600),
                  (P30, 12, 12, 12, 64, 64), (P60, 3, 3, 3, 1500, 1200)]
        items, seen = [], set()
        for (p, m, k, n, la, lb) in shapes:
            x = learn.features(m, k, n, la, lb, p)
            pred = mlp.forward(x[None, :])[0]
            valid = [c for c in range(8) if pml_amd.config_valid(ctxs[p], m, k, n, la, lb, c)]
            want = min(valid, key=lambda c: pred[c])
            got = pml_amd.choose_config(ctxs[p], m, k, n, la, lb, "auto")
            assert got == want
            seen.add(got)
            items.append((ctxs[p], O.rand_uniform((m, k, la), p, la), O.rand_uniform((k, n, lb), p, lb)))
        res = pml_amd.mul_batch(items, method="auto")
        for (ctx, A, B), C in zip(items, res):
            assert np.array_equal(C, O.polmat_mul(A, B, ctx.p))
        # the threshold rule is still available beside the model
        for (p, m, k, n, la, lb) in shapes:
            assert pml_amd.choose_config(ctxs[p], m, k, n, la, lb, "threshold") in range(8)
    finally:
        pml_amd.set_refine(True)
        learn.Mlp.uninstall()
        mlp.close()
        for c in ctxs.values():
            c.close()


REFINE_MARGIN = 0.15  # kRefineMargin, pml_amd/csrc/dispatch.hpp


def test_refinement_picks_among_close_candidates(learn):
    """With refinement on, a shape whose predicted times lie within the
    refinement margin (kRefineMargin in dispatch.hpp, 15%) of the model's best
    gets one of those candidates, the same one on every later call (cached
    per shape class), and exact products."""
    import math
    import pml_amd
    mlp = learn.Mlp.load(learn.DEFAULT_MODEL)
    ctx = pml_amd.Context(P60)
    try:
        mlp.install()
        pml_amd.set_refine(True)
        shapes = [(32, 32, 32, 4096, 4096), (24, 17, 30, 900, 1300), (7, 40, 9, 3000, 200)]
        items = []
        for (m, k, n, la, lb) in shapes:
            pred = mlp.forward(learn.features(m, k, n, la, lb, P60)[None, :])[0]
            valid = [c for c in range(8) if pml_amd.config_valid(ctx, m, k, n, la, lb, c)]
            best = min(pred[c] for c in valid)
            cand = {c for c in valid if pred[c] <= best + math.log1p(REFINE_MARGIN) + 1e-6}
            got = pml_amd.choose_config(ctx, m, k, n, la, lb, "auto")
            assert got in cand, (got, cand)
            assert pml_amd.choose_config(ctx, m, k, n, la, lb, "auto") == got
            items.append((ctx, O.rand_uniform((m, k, la), P60, la), O.rand_uniform((k, n, lb), P60, lb)))
        for (c_, A, B), C in zip(items, pml_amd.mul_batch(items, method="auto")):
            assert np.array_equal(C, O.polmat_mul(A, B, P60, method="ntt"))
    finally:
        learn.Mlp.uninstall()
        mlp.close()
        ctx.close()


def test_learned_dispatch_gains_over_threshold_on_unseen_shapes(learn):
    """The gain test: trained on the device on 20,000 recorded shapes of the
    uneven-batch workload, the network's picks on the 2,000 held-out shapes
    cost less recorded time than the threshold rule's and come close to the
    per-shape best.  (Outputs are exact under any pick: see the test above.)"""
    import pml_amd
    shapes, T = learn.load_timings()
    X, Y, M = learn.dataset(shapes[:20000], T[:20000])
    ctx_by_p = {P30: pml_amd.Context(P30), P60: pml_amd.Context(P60)}
    try:
        mlp = learn.Mlp(seed=11)
        mlp.train(X, Y, M, epochs=300, batch=256, lr=1e-3, seed=7)
        ev = learn.evaluate(mlp, shapes[20000:], T[20000:], ctx_by_p)
        print("held-out: learned %.1f ms, threshold %.1f ms, best %.1f ms, gain %.3f" % (
            ev["learned"], ev["threshold"], ev["best"], ev["gain_vs_threshold"]))
        assert ev["shapes"] == 2000
        # the baseline is the banded rule fitted on the same 20,000 shapes
        # (1.034 measured; 1.13 against the fixed three-line rule)
        assert ev["gain_vs_threshold"] > 1.015
        assert ev["learned"] <= 1.03 * ev["best"]
        mlp.close()
        packaged = learn.Mlp.load(learn.DEFAULT_MODEL)
        ev2 = learn.evaluate(packaged, shapes[20000:], T[20000:], ctx_by_p)
        assert ev2["gain_vs_threshold"] > 1.015
        packaged.close()
    finally:
        for c in ctx_by_p.values():
            c.close()


def test_banded_threshold_rule_matches_its_table(learn):
    """The library's threshold rule is the banded table (bands.inc, fitted by
    learn.fit_bands on the first 20,000 recorded shapes) with the fixed rule
    as its fallback: the library's choice equals the table lookup done here
    for 600 shapes of the workload, and refitting reproduces the table."""
    import os
    import re
    import pml_amd
    shapes, T = learn.load_timings()
    table = learn.fit_bands(shapes[:20000], T[:20000])
    src = open(os.path.join(os.path.dirname(pml_amd.__file__), "csrc", "bands.inc")).read()
    shipped = [int(x) for x in re.search(r"\{([^}]*)\}", src).group(1).split(",")]
    assert shipped == [int(x) for x in table]
    ctxs = {P60: pml_amd.Context(P60), P30: pml_amd.Context(P30)}
    try:
        for (p, m, k, n, la, lb) in learn.draw_shapes(600, 99):
            ctx = ctxs[p]
            c = int(table[learn.band_index(p, m, k, n, la, lb)])
            if not (c >= 0 and pml_amd.config_valid(ctx, m, k, n, la, lb, c)):
                if min(la, lb) <= 8 and pml_amd.config_valid(ctx, m, k, n, la, lb, 0):
                    c = 0
                elif pml_amd.config_valid(ctx, m, k, n, la, lb, 2):
                    c = 2
                else:
                    c = 1
            assert pml_amd.choose_config(ctx, m, k, n, la, lb, "threshold") == c, (p, m, k, n, la, lb)
    finally:
        for x in ctxs.values():
            x.close()
