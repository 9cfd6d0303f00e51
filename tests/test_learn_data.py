"""The packaged dispatch data: recorded timings of the uneven-batch workload
and the trained network's parameters (loaded without pickle)."""
import numpy as np

from pml_amd import learn


def test_recorded_timings_are_well_formed():
    shapes, T = learn.load_timings()
    assert len(shapes) == 22000 and T.shape == (22000, 8)
    for (p, m, k, n, la, lb) in shapes[:2000]:
        assert p in (learn.P30, learn.P60)
        assert 2 <= m <= 64 and 2 <= k <= 64 and 2 <= n <= 64
        assert 16 <= la <= 16384 and 16 <= lb <= 16384
    # every shape has at least the plain transform configurations timed
    assert (T[:, 1] > 0).all() and (T[:, 3] > 0).all() and (T[:, 7] > 0).all()
    # untimed entries are marked -1, timed ones are positive milliseconds
    assert ((T == -1) | (T > 0)).all()


def test_default_model_parameters_load():
    with np.load(learn.DEFAULT_MODEL, allow_pickle=False) as z:
        theta = z["params"]
    assert theta.shape == (9224,) and theta.dtype == np.float32 and np.isfinite(theta).all()
