"""Build libpmlgpu.so (HIP kernels for gfx950 + C host layer) in place.

Usage: python -m pml_amd.build   (or pml_amd.build.build() from Python)

Every HIP translation unit is compiled with hipcc for gfx950 only; the
FLINT-shaped C drop-in layer is compiled with gcc; both are linked into
pml_amd/libpmlgpu.so next to this file.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpmlgpu.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIP_SOURCES = ["ntt.hip", "pointwise.hip", "api.hip", "crt.hip", "pmbasis.hip", "kernel.hip",
               "staging.hip", "batch.hip", "mlp.hip"]
C_SOURCES = ["dropin.c"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# per-file extra flags: keep MFMA results in VGPRs (no AGPR round trips before the fold)
HIP_EXTRA = {"pointwise.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", INCLUDE,
]
C_FLAGS = ["-O2", "-fPIC", "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-I", INCLUDE]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # the host pass of hipcc does not know the packed-fp32 feature name; harmless
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + HIP_EXTRA.get(src, []) + ["-c", path, "-o", obj])
    for src in C_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD, src + ".o")
        jobs.append(["gcc"] + C_FLAGS + ["-c", path, "-o", obj])
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
