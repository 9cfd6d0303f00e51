"""Full-size GPU results against the oracle's golden digests
(tests/golden/golden.json, written by tests/golden/gen_golden.py): the
32x32x32 length-4096 product with both pointwise kernels, the PM-Basis
middle product at its top-level shape, and the 64x32 order-2048 PM-Basis."""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden.json")))


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def pml():
    import pml_amd
    return pml_amd


@pytest.mark.parametrize("key", ["c1_mul_4x4x4_len256_p60", "c2_mul_32x32x32_len4096_p60"])
@pytest.mark.parametrize("method", ["mfma", "plain"])
def test_products(pml, key, method):
    g = GOLD[key]
    A = O.rand_uniform((g["m"], g["k"], g["lenA"]), g["p"], g["seedA"])
    B = O.rand_uniform((g["k"], g["n"], g["lenB"]), g["p"], g["seedB"])
    with pml.Context(g["p"]) as ctx:
        C = ctx.mul(A, B, method=method)
    assert [int(x) for x in C[0, 0, :8]] == g["C000_first8"]
    assert digest(C) == g["sha256"]


@pytest.mark.parametrize("method", ["mfma", "plain"])
def test_mulmid_top_level(pml, method):
    g = GOLD["mulmid_64x64x32_len1025x2048_1024_2048_p30"]
    A = O.rand_uniform((g["m"], g["k"], g["lenA"]), g["p"], g["seedA"])
    B = O.rand_uniform((g["k"], g["n"], g["lenB"]), g["p"], g["seedB"])
    with pml.Context(g["p"]) as ctx:
        R = ctx.mulmid(A, B, g["nlo"], g["nhi"], method=method)
    assert digest(R) == g["sha256"]


@pytest.mark.parametrize("key", ["c3_pmbasis_64x32_order2048_p30", "pmbasis_12x5_order300_p60_shift3i"])
def test_pmbasis(pml, key):
    g = GOLD[key]
    F = O.rand_uniform((g["m"], g["n"], g["lenF"]), g["p"], g["seedF"])
    shift = np.zeros(g["m"], dtype=np.int64) if g["shift_in"] == 0 else np.array(g["shift_in"], dtype=np.int64)
    with pml.Context(g["p"]) as ctx:
        P, sh = ctx.pmbasis(F, g["order"], shift)
    P = O.normalise(P)
    assert P.shape[2] == g["len"]
    assert [int(x) for x in sh] == g["shift_out"]
    assert digest(P) == g["sha256"]
