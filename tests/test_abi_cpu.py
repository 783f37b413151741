"""C-ABI library checks that need no GPU: it is built, loads, and exports
every entry point include/fgx.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fgx.h")
LIB = os.path.join(ROOT, "beofaunus_amd", "_lib", "libfgx.so")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fgx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_core_entry_points():
    syms = declared_symbols()
    for s in ("fgx_create", "fgx_energy_change", "fgx_commit", "fgx_propagate",
              "fgx_propagate_batch", "fgx_system_energy", "fgx_g2g", "fgx_i2all",
              "fgx_temper_exchange"):
        assert s in syms


def test_library_loads_and_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail("libfgx.so not built; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_binding_signatures_cover_header():
    from beofaunus_amd import fgx
    L = fgx.lib()
    for s in declared_symbols():
        assert getattr(L, s).restype is not None or s == "fgx_destroy"
