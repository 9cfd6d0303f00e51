"""Host-side pieces of bench.py, on the CPU: the generator that regenerates
the golden fixture's inputs without the oracle (the PM-Basis leg compares its
basis with the fixture's digest) matches the oracle's seeded stream."""
import numpy as np

import bench
from oracle import oracle as O


def test_bench_generator_matches_oracle_stream():
    for p, seed, n in [(998244353, 401, 5000), (1152921504556515329, 201, 3000), (7, 3, 500),
                       (2**63 + 29, 9, 100)]:
        assert np.array_equal(bench.splitmix_uniform(n, p, seed), O.rand_uniform((n,), p, seed))
