"""The C-ABI library loads on a machine without a GPU and exports every
function that include/*.h declares.  No compute call is made here."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "pml_amd", "libpmlgpu.so")

DECL = re.compile(r"^[A-Za-z_][\w \*]*?\b((?:pmlg|nmod|_nmod)\w*)\s*\(", re.M)


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for mm in DECL.finditer(text):
            names.add(mm.group(1))
    return sorted(names)


def test_headers_declare_the_entry_points():
    names = declared_functions()
    for must in ("pmlg_ctx_create", "pmlg_polmat_mul", "pmlg_polmat_mul_dev",
                 "pmlg_polmat_mulmid", "pmlg_polmat_mulmid_dev"):
        assert must in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libpmlgpu.so is not built (python -m pml_amd.build)")
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_front_end_binds():
    import pml_amd
    L = pml_amd.lib()
    assert L.pmlg_last_error() is not None
