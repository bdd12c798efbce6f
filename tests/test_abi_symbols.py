"""The C-ABI library loads and exports every symbol include/pfgpu.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "pfgpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pfgpu_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_solver_entry_points():
    syms = declared_symbols()
    for s in ("pfgpu_ctx_create", "pfgpu_grid_load", "pfgpu_solve_batch", "pfgpu_solve_batch_device"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from lightsim2grid_amd import abi
    L = ctypes.CDLL(abi.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(L, s), s
    assert set(declared_symbols()) == set(abi.SIGNATURES)
    assert L.pfgpu_abi_version() == 2
