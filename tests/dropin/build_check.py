"""Compiles tests/dropin/dropin_check (the drop-in parity program) and
tests/dropin/dropin_bench (its timing program) with gcc,
linked against pml_amd/libpmlgpu.so and the oracle's liboracle.so.

    python tests/dropin/build_check.py
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "dropin_check.c")
EXE = os.path.join(HERE, "dropin_check")
BENCH_SRC = os.path.join(HERE, "dropin_bench.c")
BENCH_EXE = os.path.join(HERE, "dropin_bench")


def _compile(src, exe, opt):
    lib = os.path.join(ROOT, "pml_amd")
    orc = os.path.join(ROOT, "oracle")
    cmd = ["gcc", opt, "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-I", os.path.join(ROOT, "include"), "-I", orc, src, "-o", exe,
           "-L", lib, "-lpmlgpu", "-L", orc, "-loracle",
           # relative run paths: the tree may be copied elsewhere
           "-Wl,-rpath,$ORIGIN/../../pml_amd", "-Wl,-rpath,$ORIGIN/../../oracle"]
    subprocess.run(cmd, check=True)


def build():
    """The parity program and the drop-in timing program (dropin_bench)."""
    _compile(SRC, EXE, "-O1")
    _compile(BENCH_SRC, BENCH_EXE, "-O2")
    return EXE


if __name__ == "__main__":
    print(build())
