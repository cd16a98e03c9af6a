"""The C-ABI library: loads, exports every declared entry point, reports
errors through status codes.  CPU only (no compute on a device)."""
import ctypes as C
import os
import re
import subprocess

import pytest

import fvmhip as F
from fvmhip import _lib as L
from tests.conftest import ROOT, has_gpu

HEADER = os.path.join(ROOT, "include", "fvmhip.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(fvh_\w+)\s*\(", txt, re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = L.lib()
    names = declared()
    assert len(names) >= 45
    for n in names:
        assert hasattr(lib, n), f"{n} declared in fvmhip.h but not exported"
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (fvh_\w+)", out))
    assert set(names) <= exported
    assert set(n for n in exported) - set(names) == set(), "undeclared exports"


def test_bindings_cover_header():
    assert set(declared()) == set(L.SIGNATURES)


def test_version_and_status_strings():
    assert b"gfx950" in L.lib().fvh_version()
    assert L.lib().fvh_status_string(0) == b"ok"
    assert b"mesh" in L.lib().fvh_status_string(L.FVH_ERR_MESH)


def test_bad_arguments_are_status_codes_not_crashes():
    lib = L.lib()
    assert lib.fvh_hmesh_read_unv(b"/nonexistent.unv", 1.0, C.byref(C.c_void_p())) == L.FVH_ERR_IO
    buf = C.create_string_buffer(256)
    lib.fvh_last_error(buf, 256)
    assert b"nonexistent" in buf.value
    assert lib.fvh_mesh_create(None, C.byref(C.c_void_p())) == L.FVH_ERR_ARG
    assert lib.fvh_spmv(None, None, None) == L.FVH_ERR_ARG
    with pytest.raises(F.FVHError):
        F.build_mesh("/nonexistent.unv")


@pytest.mark.skipif(has_gpu(), reason="checks the no-device error path")
def test_device_calls_fail_loudly_without_a_gpu(quad40):
    assert F.device_count() == 0
    d = quad40.desc()
    assert L.lib().fvh_mesh_create(C.byref(d), C.byref(C.c_void_p())) == L.FVH_ERR_NODEV
    with pytest.raises(F.FVHError):
        F.Simulation(__import__("tests.cases").cases.cavity_model(quad40), __import__("tests.cases").cases.config(1))


def test_product_does_not_use_the_oracle():
    pkg = os.path.join(ROOT, "xcalibre_tw.jl_amd")
    for dp, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".h")):
                txt = open(os.path.join(dp, f)).read()
                assert "import oracle" not in txt and "from oracle" not in txt and "ref_cpu" not in txt, f
