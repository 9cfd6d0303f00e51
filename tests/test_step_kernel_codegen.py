"""Code generation of the M-Basis step kernels (CPU: hipcc cross-compiles).

The step kernel runs in one workgroup and is latency-bound, so what the
compiler makes of its lambdas matters: when the composition was left as a
function call in the two-row kernels, its captures and arrays went through
scratch memory and a 128x96 basis took 2.2 times as long.  This compiles
pmbasis.hip to gfx950 assembly with the library's flags and checks that no
step-kernel instantiation calls a function or uses scratch memory."""
import os
import re
import shutil
import subprocess

import pytest

from pml_amd import build

HIPCC = build.HIPCC


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not installed")
def test_step_kernels_inline_and_without_scratch(tmp_path):
    out = tmp_path / "pmbasis.s"
    cmd = [HIPCC] + build.HIP_FLAGS + ["--cuda-device-only", "-S",
                                       os.path.join(build.CSRC, "pmbasis.hip"), "-o", str(out)]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    asm = out.read_text()
    # kernel bodies: from each step kernel's label to its end marker
    bodies = re.findall(r"^(_ZN4pmlg\w*mbasis_step_kernel\w*):.*?$(.*?)^\.Lfunc_end", asm, re.S | re.M)
    assert len(bodies) == 6  # {Mont64, Mont32, Big} x {one, two rows per lane}
    for name, body in bodies:
        assert "s_swappc" not in body, name
    # kernel descriptors: scratch bytes per lane
    descs = re.findall(r"\.amdhsa_kernel (_ZN4pmlg\w*mbasis_step_kernel\w*)\s.*?"
                       r"\.amdhsa_private_segment_fixed_size (\d+)", asm, re.S)
    assert len(descs) == 6
    for name, scratch in descs:
        assert int(scratch) == 0, (name, scratch)
