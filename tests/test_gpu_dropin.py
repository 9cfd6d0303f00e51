"""The FLINT-shaped drop-in (include/pmlgpu_flint.h), called from C the way
a PML user calls nmod_poly_mat_multiply / mulmid / pmbasis / mbasis /
is_approximant_basis, checked against the oracle by tests/dropin/dropin_check.c.
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "dropin", "dropin_check")


@pytest.mark.gpu
def test_dropin_entry_points_match_oracle():
    if not os.path.exists(EXE):
        from tests.dropin import build_check
        build_check.build()
    r = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=110)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-3000:]
    assert "all drop-in checks passed" in r.stdout
