"""CPU-side checks of the C ABI: the library loads and exports every entry
point include/pdgpu.h declares (no compute calls: no GPU here)."""
import os
import re

import pytest

from peridigm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    out = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            txt = open(os.path.join(inc, fn)).read()
            out |= set(re.findall(r"^\s*(?:int|const char\s*\*)\s+(pdg_\w+)\s*\(", txt, re.M))
    return out


def test_library_built():
    assert os.path.exists(_lib.LIB_PATH), "libpdgpu.so missing: run __graft_entry__.build()"


def test_every_declared_symbol_exported_and_bound():
    decl = declared()
    assert len(decl) >= 20
    assert decl <= set(_lib.SIGNATURES), decl - set(_lib.SIGNATURES)
    assert decl <= _lib.exported_symbols()


def test_library_loads_without_gpu():
    lib = _lib.load()
    assert hasattr(lib, "pdg_ctx_create")
