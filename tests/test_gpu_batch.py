"""GPU parity of the batched products (pmlg_batch_plan_*).

Every product of a batch is compared bit for bit with the CPU oracle, under
every configuration the dispatcher can pick (direct convolution, transforms
with the plain or MFMA pointwise kernel, other transform radices and MFMA
block sizes), with and without HIP-graph replay.  Batches mix both moduli of
the uneven-batch workload (a 30-bit and a 60-bit FFT prime), uneven shapes and
lengths, transform lengths above the LDS limit, and the empty cases.  The
reference has no batched entry point; its own tests check single products
against FLINT (flint-extras/src/nmod_poly_mat_extra/test/t-mul_geometric.c);
the oracle is that check here.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

P60 = 1152921504556515329
P30 = 998244353


@pytest.fixture(scope="module")
def pml():
    import pml_amd
    return pml_amd


@pytest.fixture(scope="module")
def ctxs(pml):
    c = {P60: pml.Context(P60), P30: pml.Context(P30)}
    yield c
    for x in c.values():
        x.close()


def draw_batch(count, seed, max_dim=12, log_len=(4, 10), big=()):
    """Shapes of an uneven batch: dims uniform, lengths log-uniform, primes 50/50."""
    rng = np.random.default_rng(seed)
    shapes = []
    for i in range(count):
        p = P30 if rng.integers(0, 2) == 0 else P60
        m, k, n = (int(x) for x in rng.integers(2, max_dim + 1, size=3))
        la = int(2 ** rng.uniform(*log_len))
        lb = int(2 ** rng.uniform(*log_len))
        shapes.append((p, m, k, n, la, lb))
    shapes.extend(big)
    return shapes


def inputs(shapes, seed):
    out = []
    for i, (p, m, k, n, la, lb) in enumerate(shapes):
        A = O.rand_uniform((m, k, la), p, seed + 2 * i)
        B = O.rand_uniform((k, n, lb), p, seed + 2 * i + 1)
        out.append((p, A, B))
    return out


def expected(data):
    return [O.polmat_mul(A, B, p) for (p, A, B) in data]


BIG = [
    # transform lengths above the LDS limit (radix split + merge) for both word sizes
    (P60, 2, 3, 2, 9000, 7000),   # N = 2^14 on 64-bit words
    (P60, 2, 2, 1, 12000, 9000),  # N = 2^15
    (P30, 3, 2, 2, 12000, 9000),  # N = 2^15 on 32-bit words
    (P30, 2, 2, 2, 8000, 300),    # N = 2^13
]


@pytest.fixture(scope="module")
def batch_data():
    shapes = draw_batch(40, 7, big=BIG)
    data = inputs(shapes, 1000)
    return shapes, data, expected(data)


def run_batch(pml, ctxs, data, **kw):
    items = [(ctxs[p], A, B) for (p, A, B) in data]
    return pml.mul_batch(items, **kw)


@pytest.mark.parametrize("method", ["auto", "threshold", "plain", "mfma"])
def test_batch_methods_match_oracle(pml, ctxs, batch_data, method):
    shapes, data, ref = batch_data
    got = run_batch(pml, ctxs, data, method=method)
    for i, (g, r) in enumerate(zip(got, ref)):
        assert np.array_equal(g, r), "product %d %s differs (method %s)" % (i, shapes[i], method)


def test_batch_every_config_matches_oracle(pml, ctxs, batch_data):
    shapes, data, ref = batch_data
    for c in range(pml.num_configs()):
        sel = [i for i, (p, m, k, n, la, lb) in enumerate(shapes)
               if pml.config_valid(ctxs[p], m, k, n, la, lb, c)]
        if c == 0:  # direct convolution: keep the timed cost small
            sel = [i for i in sel if shapes[i][4] * shapes[i][5] <= 1 << 16]
        assert sel, "config %d offered for no product" % c
        got = run_batch(pml, ctxs, [data[i] for i in sel], config=c)
        for j, i in enumerate(sel):
            assert np.array_equal(got[j], ref[i]), "config %d, product %d %s" % (c, i, shapes[i])


def test_batch_graph_replays(pml, ctxs, batch_data):
    """A captured plan replayed several times, with the outputs cleared and
    other device work between replays, reproduces the eager result each time."""
    shapes, data, ref = batch_data
    items = [(ctxs[p], A, B) for (p, A, B) in data[:24]]
    with pml.BatchPlan(items, graph=True) as plan:
        launches, arena = plan.stats()
        assert launches < 3 * len(items), "no grouping: %d launches" % launches
        for rep in range(3):
            for (ctx, dims, dC) in plan.items:  # clear outputs
                m, k, n, la, lb, lc = dims
                dC.upload(np.zeros(max(m * n * lc, 1), dtype=np.uint64))
            # unrelated device work on a context stream between replays
            ctxs[P60].mul(O.rand_uniform((2, 2, 64), P60, rep), O.rand_uniform((2, 2, 64), P60, rep + 9))
            plan.run()
            plan.sync()
            for i, g in enumerate(plan.results()):
                assert np.array_equal(g, ref[i]), "replay %d, product %d" % (rep, i)


def test_batch_stage_timing_replays(pml, ctxs, batch_data):
    """Stage timing of replays: one graph per stage, HIP events between them;
    every replay still gives the oracle's products, and the stage times of
    the replays add up to at most their wall time."""
    import time
    shapes, data, ref = batch_data
    items = [(ctxs[p], A, B) for (p, A, B) in data[:16]]
    with pml.BatchPlan(items, graph=True) as plan:
        plan.time_stages(4)
        plan.run()  # eager pass and capture
        plan.sync()
        plan.stage_times()
        t0 = time.perf_counter()
        for _ in range(6):
            plan.run()
        plan.sync()
        wall_ms = (time.perf_counter() - t0) * 1e3
        ms, n = plan.stage_times()
        assert n == 4 and all(x > 0 for x in ms)
        assert sum(ms) * 6 <= wall_ms * 1.05 + 1.0, (ms, wall_ms)
        for i, g in enumerate(plan.results()):
            assert np.array_equal(g, ref[i]), "product %d" % i


def test_batch_empty_and_degenerate(pml, ctxs):
    p = P60
    rng_items = [
        (p, O.rand_uniform((3, 0, 5), p, 1), O.rand_uniform((0, 4, 6), p, 2)),  # k = 0: zero output
        (p, O.rand_uniform((2, 2, 0), p, 3), O.rand_uniform((2, 3, 4), p, 4)),  # lenA = 0: no output
        (p, O.rand_uniform((1, 1, 1), p, 5), O.rand_uniform((1, 1, 1), p, 6)),  # constants
        (P30, O.rand_uniform((2, 5, 1), P30, 7), O.rand_uniform((5, 3, 300), P30, 8)),
    ]
    got = run_batch(pml, ctxs, rng_items)
    assert got[0].shape == (3, 4, 10) and not got[0].any()
    assert got[1].shape == (2, 3, 0)
    for i in (2, 3):
        p_, A, B = rng_items[i]
        assert np.array_equal(got[i], O.polmat_mul(A, B, p_))


def test_batch_extreme_residues(pml, ctxs):
    """All coefficients p-1: the largest dot products each kernel must fold."""
    items = []
    for p in (P60, P30):
        for (m, k, n, la, lb) in [(4, 64, 4, 40, 50), (3, 64, 2, 3, 5), (2, 30, 2, 700, 600)]:
            A = np.full((m, k, la), p - 1, dtype=np.uint64)
            B = np.full((k, n, lb), p - 1, dtype=np.uint64)
            items.append((p, A, B))
    for c in range(pml.num_configs()):
        sel = [(p, A, B) for (p, A, B) in items
               if pml.config_valid(ctxs[p], A.shape[0], A.shape[1], B.shape[1], A.shape[2],
                                   B.shape[2], c)]
        got = run_batch(pml, ctxs, sel, config=c)
        for (p, A, B), g in zip(sel, got):
            assert np.array_equal(g, O.polmat_mul(A, B, p)), "config %d shape %s" % (c, A.shape)


def test_autotune_records_best(pml, ctxs):
    ctx = ctxs[P60]
    ms, best = pml.autotune(ctx, 8, 8, 8, 300, 200, reps=2)
    timed = [c for c, t in enumerate(ms) if t >= 0]
    assert best in timed and ms[best] == min(ms[c] for c in timed)
    assert 0 in timed and 2 in timed  # direct convolution and MFMA both timed at this size
    assert pml.choose_config(ctx, 8, 8, 8, 300, 200, "auto") == best
    assert pml.choose_config(ctx, 8, 8, 8, 300, 200, "threshold") in timed
