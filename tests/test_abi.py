"""The C ABI without a GPU: both libraries load, every function that
include/mxc_gpu.h and include/mxc_h2.h declare is exported by the library
that implements it, the Python bindings name only declared functions, and
entry points reject null handles with an error code instead of touching the
device."""
import ctypes as C
import os
import re

import pytest

from mxc_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    """Function names declared in a header (declarations may span lines)."""
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    text = re.sub(r"//[^\n]*", " ", text)
    names = re.findall(r"\b(mx[cg]_[A-Za-z0-9_]+)\s*\(", text)
    # typedef'd function-pointer or struct names are not functions
    return sorted(set(n for n in names if not n.endswith("_t")))


@pytest.mark.parametrize("header,lib", [("mxc_gpu.h", "gpu"), ("mxc_h2.h", "h2")])
def test_every_declared_function_is_exported(header, lib):
    L = N.gpu() if lib == "gpu" else N.h2()
    names = declared(header)
    assert len(names) > 10, names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_name_declared_functions():
    src = open(os.path.join(ROOT, "mxc_amd", "_native.py")).read()
    bound = set(re.findall(r'"(mx[cg]_[A-Za-z0-9_]+)"\s*:', src))
    decl = set(declared("mxc_gpu.h")) | set(declared("mxc_h2.h"))
    assert bound and bound <= decl, sorted(bound - decl)


def test_null_handles_are_rejected_without_a_device():
    G = N.gpu()
    assert G.mxg_h2_matvec(None, None, None, 1, 1, N.MXG_FP64) != 0
    assert G.mxg_ulv_solve(None, None, 1, 1) != 0
    assert b"null" in G.mxg_last_error().lower()
    w = C.c_double()
    assert G.mxg_profile_work(None, 0, C.byref(w)) != 0
    runs = (C.c_longlong * 2)(0, 1)
    assert G.mxg_h2_construct_level_runs(None, 0, 1, None, None, None, None, None, None, None, runs, 0.0, 0, 0,
                                         None) != 0
