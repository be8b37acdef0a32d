"""The C-ABI library loads and exports every symbol include/exafmm_hip.h declares.

No compute call is made here (no GPU in the CPU tier)."""
import ctypes
import os
import re

import exafmm_hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            syms |= set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(exafmm_hip_\w+)\s*\(", text, re.M))
    return syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(exafmm_hip.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in sorted(syms) if not hasattr(lib, s)]
    assert not missing, missing
    assert set(exafmm_hip.EXPORTS) == syms


def test_binding_loads():
    L = exafmm_hip.load()
    assert L.exafmm_hip_create is not None
