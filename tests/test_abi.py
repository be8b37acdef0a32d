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


def test_md_wrapper_exports_and_header_compiles(tmp_path):
    """The MD-code entry points (exafmm_gromacs.h) are exported, and the
    header compiles as C and as C++ (int & n under C++, int *n under C)."""
    import subprocess
    text = open(os.path.join(ROOT, "include", "exafmm_gromacs.h")).read()
    syms = set(re.findall(r"^\s*void\s+(\w+)\s*\(", text, re.M))
    assert syms == {"FMM_Init", "FMM_Finalize", "FMM_Partition", "FMM_Coulomb", "Ewald_Coulomb", "Direct_Coulomb"}
    lib = ctypes.CDLL(exafmm_hip.LIB_PATH)
    assert all(hasattr(lib, s) for s in syms)
    for cc, lang in (("gcc", "c"), ("g++", "c++")):
        src = tmp_path / f"use.{lang.replace('+', 'p')}"
        src.write_text('#include "exafmm_gromacs.h"\nint main(void) { FMM_Init(0); FMM_Finalize(); return 0; }\n')
        subprocess.run([cc, "-x", lang, "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        str(src)], check=True)
