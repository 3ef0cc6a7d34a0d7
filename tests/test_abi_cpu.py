"""CPU-side checks of the C ABI: the in-tree library loads, exports every
entry point declared in include/*.h, and refuses to run without a gfx950
device (there is no CPU fallback).  No kernel is launched here."""
import ctypes
import glob
import os
import re

import pytest

from tests.conftest import ROOT


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(akx_\w+)\s*\(", src))
    return sorted(names)


def test_header_declares_entry_points():
    names = _declared()
    assert "akx_model_solve_step" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    from akantu2_amd import _native
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from akantu2_amd import _native
    lib = _native.lib()
    for n in _declared():
        assert getattr(lib, n).restype is not None or n.startswith("akx_")


def test_no_gpu_means_no_model():
    from akantu2_amd import _native, box_mesh, SolidMechanicsModel
    if _native.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="no gfx950 device"):
        SolidMechanicsModel(box_mesh("_tetrahedron_4", 2, 1, 1))


def test_law_lookup_by_reference_name():
    from akantu2_amd import _native as N
    lib = N.lib()
    law = ctypes.c_int(-1)
    N.check(lib.akx_material_law_by_name(b"neohookean", 0, ctypes.byref(law)))
    assert law.value == N.LAW_NEOHOOKEAN
    N.check(lib.akx_material_law_by_name(b"elastic", 1, ctypes.byref(law)))
    assert law.value == N.LAW_ELASTIC_FD
    with pytest.raises(N.AkantuError, match="unknown material law"):
        N.check(lib.akx_material_law_by_name(b"mazars", 0, ctypes.byref(law)))


def test_product_does_not_import_oracle():
    """The product package never loads the oracle (test infrastructure)."""
    for path in glob.glob(os.path.join(ROOT, "akantu2_amd", "**", "*.py"),
                          recursive=True):
        src = open(path).read()
        assert not re.search(r"^\s*(import|from)\s+oracle", src, re.M), path


def test_cpp_host_api_compiles():
    """include/akantu_mi355x.hpp (the C++ API above the C ABI) compiles, and
    the C++ driver used by the GPU test is built."""
    import subprocess
    src = os.path.join(ROOT, "tests", "cpp", "drive_explicit.cc")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall",
                        "-I" + os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
