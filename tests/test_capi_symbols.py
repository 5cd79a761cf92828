"""The C-ABI library loads on the CPU host and exports every entry point that
include/vfvm_hip.h declares (no compute call: no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "voronoifvm.jl_amd", "libvfvm_hip.so")


def declared_functions():
    text = open(os.path.join(ROOT, "include", "vfvm_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vfvm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("vfvm_ctx_create", "vfvm_system_create", "vfvm_assemble", "vfvm_linsolve", "vfvm_newton",
                 "vfvm_pattern"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_covers_header():
    import vfvm._capi as capi
    assert set(declared_functions()) <= set(capi.SIGNATURES)
