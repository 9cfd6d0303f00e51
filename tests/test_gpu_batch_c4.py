"""The uneven-batch workload at its full size, checked against the oracle.

512 independent products A_i (m x k) * B_i (k x n) with m, k, n uniform in
2..64, lengths log-uniform in 16..16384 and the 30-bit / 60-bit FFT primes
drawn 50/50 (pml_amd.learn.draw_shapes, seed 4: the batch that
`python -m pml_amd.perf batch` times), operands generated on the device, run
as one batch plan under the learned dispatch with the packaged network and
replayed as a HIP graph.  The reference has no batched entry point; its own
tests compare each product with a second method
(flint-extras/src/nmod_poly_mat_extra/test/t-mul_geometric.c:24-80).  Here,
after two replays (the outputs overwritten with other residues before the
second):

* every product satisfies C(r) = A(r) B(r) mod p at two seeded points, the
  evaluations by the oracle's Horner scheme and A(r) B(r) by the oracle's
  closed-form product;
* 32 products are compared in full with the oracle's NTT product: the 8 with
  the longest output, the 8 with the largest m k n, and 16 more drawn from a
  seeded generator.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEED = 4


@pytest.fixture(scope="module")
def c4():
    from pml_amd import learn, lib, _check
    from pml_amd.perf import DeviceBatch
    model = learn.install_default()
    shapes = learn.draw_shapes(512, SEED)
    batch = DeviceBatch(shapes, "auto")
    plan = batch.plan(1)  # graph capture on the first run
    try:
        _check(lib().pmlg_batch_plan_run(plan, None))
        _check(lib().pmlg_batch_plan_sync(plan))
        batch.scramble_outputs(777)
        _check(lib().pmlg_batch_plan_run(plan, None))  # graph replay
        _check(lib().pmlg_batch_plan_sync(plan))
        print("C4 batch: 512 products planned and replayed", flush=True)
        yield batch
    finally:
        lib().pmlg_batch_plan_destroy(plan)
        batch.close()
        model.close()


def test_c4_shapes_cover_the_workload(c4):
    sh = np.array([s[1:] for s in c4.shapes])
    assert len(c4.shapes) == 512
    assert sh[:, :3].min() == 2 and sh[:, :3].max() == 64
    assert sh[:, 3:].min() >= 16 and sh[:, 3:].max() >= 12000
    ps = {s[0] for s in c4.shapes}
    assert ps == {998244353, 1152921504556515329}


def test_c4_every_product_evaluation_identity(c4):
    rng = np.random.default_rng(SEED)
    bad = []
    for i, (p, m, k, n, la, lb) in enumerate(c4.shapes):
        A, B, C = c4.operands(i)
        for r in (int(x) for x in rng.integers(2, p - 1, size=2)):
            Ar = O.polmat_eval(A, r, p)[:, :, None]
            Br = O.polmat_eval(B, r, p)[:, :, None]
            Cr = O.polmat_eval(C, r, p)
            if not np.array_equal(O.polmat_mul(Ar, Br, p, "school")[:, :, 0], Cr):
                bad.append((i, c4.shapes[i], r))
        if (i + 1) % 128 == 0:
            print("evaluation identity: %d/512 products checked" % (i + 1), flush=True)
    assert not bad, bad[:5]


def test_c4_products_in_full(c4):
    shapes = c4.shapes
    idx = list(range(len(shapes)))
    longest = sorted(idx, key=lambda i: -(shapes[i][4] + shapes[i][5]))[:8]
    largest = [i for i in sorted(idx, key=lambda i: -(shapes[i][1] * shapes[i][2] * shapes[i][3]))
               if i not in longest][:8]
    rest = [i for i in idx if i not in longest and i not in largest]
    extra = [int(i) for i in np.random.default_rng(SEED + 1).choice(rest, size=16, replace=False)]
    sel = longest + largest + extra
    assert len(set(sel)) == 32
    for t, i in enumerate(sel):
        p = shapes[i][0]
        A, B, C = c4.operands(i)
        assert np.array_equal(C, O.polmat_mul(A, B, p, "ntt")), "product %d %s" % (i, shapes[i])
        if (t + 1) % 8 == 0:
            print("full comparison: %d/32 products" % (t + 1), flush=True)
