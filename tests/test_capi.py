"""CPU checks of the C ABI: the library loads and exports every symbol that
include/*.h declares, and the Python binding covers exactly those symbols.
No compute calls (no GPU here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = {}
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^(?!\s*typedef)\s*[A-Za-z_][\w \*]*?\b([a-z_][a-z0-9_]*)\s*\(", text,
                             flags=re.M):
            name = m.group(1)
            if name in ("if", "for", "while", "return", "sizeof", "defined"):
                continue
            names[name] = h
    return names


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "alp_hip_mxv" in names and "alp_hip_cg" in names


def test_library_exports_every_declared_symbol():
    import alp_amd
    if not os.path.exists(alp_amd.LIB_PATH):
        pytest.skip("libalp_hip.so not built")
    lib = ctypes.CDLL(alp_amd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_no_undeclared_exports():
    import alp_amd
    if not os.path.exists(alp_amd.LIB_PATH):
        pytest.skip("libalp_hip.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", alp_amd.LIB_PATH], capture_output=True, text=True)
    exported = {l.split()[-1] for l in out.stdout.splitlines() if " T " in l}
    ours = {e for e in exported if e.startswith(("alp_hip_", "sparse_", "spblas_"))}
    declared = set(declared_functions())
    assert ours <= declared, ours - declared


def test_python_binding_matches_header():
    import alp_amd
    declared = {n for n, h in declared_functions().items() if h == "alp_hip.h"}
    assert set(alp_amd.SIGNATURES) == declared


def test_rc_values_match_grb_rc():
    # include/graphblas/rc.hpp:47-155 in ALP: SUCCESS = 0, then PANIC .. FAILED in order
    text = open(os.path.join(INCLUDE, "alp_hip.h")).read()
    order = ["SUCCESS", "PANIC", "OUTOFMEM", "MISMATCH", "OVERLAP", "OVERFLW", "UNSUPPORTED",
             "ILLEGAL", "FAILED"]
    for i, name in enumerate(order):
        assert re.search(rf"ALP_{name} = {i},?", text), name
