"""The C-ABI library loads on a CPU-only host and exports every symbol that
include/flame3d.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "flame3d.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(f3d_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    for must in ("f3d_create", "f3d_flux_split", "f3d_weno", "f3d_viscous_flux",
                 "f3d_spec_viscous_flux", "f3d_div", "f3d_div_specs", "f3d_mixture",
                 "f3d_c2prim", "f3d_rk3_stage", "f3d_chem_stiff", "f3d_chem_newton"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import flame3d_amd.abi as abi
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("libflame3d.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(abi.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for n in declared_symbols():
        assert re.search(rf"\bT {n}$", out, re.M), n
    # the Python binding declares exactly the header's entry points
    assert sorted(abi.exported_symbols()) == declared_symbols()


def test_binding_loads_and_reports_errors_without_gpu():
    import flame3d_amd.abi as abi
    if not os.path.exists(abi.LIB_PATH):
        pytest.skip("libflame3d.so not built")
    lib = abi.load_library()
    # a null context is rejected on the host before any HIP call
    rc = lib.f3d_flux_split(None, None, None, None, None, None, None, None, None)
    assert rc == -1
    assert b"null" in lib.f3d_last_error()
