"""Compiles tests/dropin/dropin_check (the drop-in parity program) with gcc,
linked against pml_amd/libpmlgpu.so and the oracle's liboracle.so.

    python tests/dropin/build_check.py
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "dropin_check.c")
EXE = os.path.join(HERE, "dropin_check")


def build():
    lib = os.path.join(ROOT, "pml_amd")
    orc = os.path.join(ROOT, "oracle")
    cmd = ["gcc", "-O1", "-std=gnu11", "-Wall", "-Wextra", "-Wno-unused-parameter",
           "-I", os.path.join(ROOT, "include"), "-I", orc, SRC, "-o", EXE,
           "-L", lib, "-lpmlgpu", "-L", orc, "-loracle",
           # relative run paths: the tree may be copied elsewhere
           "-Wl,-rpath,$ORIGIN/../../pml_amd", "-Wl,-rpath,$ORIGIN/../../oracle"]
    subprocess.run(cmd, check=True)
    return EXE


if __name__ == "__main__":
    print(build())
