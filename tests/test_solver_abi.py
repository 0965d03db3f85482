"""Host-side behaviour of the transition CG C ABI (alp_hip_solver.h; ALP's
include/transition/solver.h and src/transition/solver.cpp): argument checks that
return before any device work, so they run without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def S():
    import alp_amd
    if not os.path.exists(alp_amd.LIB_PATH):
        pytest.skip("libalp_hip.so not built")
    from alp_amd import solver
    return solver


def test_error_enum_matches_reference_order():
    # include/transition/solver.h:120-155: NO_ERROR .. UNKNOWN = 0 .. 5
    text = open(os.path.join(ROOT, "include", "alp_hip_solver.h")).read()
    body = re.search(r"typedef enum \{(.*?)\} sparse_err_t;", text, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = [t.strip() for t in body.split(",") if t.strip()]
    assert names == ["NO_ERROR", "NULL_ARGUMENT", "ILLEGAL_ARGUMENT", "OUT_OF_MEMORY", "FAILED", "UNKNOWN"]


@pytest.mark.parametrize("variant", ["sii", "siz", "szz", "dii", "diz", "dzz"])
def test_init_argument_checks(S, variant):
    dt, jt, it = S.VARIANTS[variant]
    a = np.ones(1, dt)
    ja = np.zeros(1, jt)
    ia = np.array([0, 1], it)
    h = ctypes.c_void_p()
    init = S.fn("init", variant)
    # solver.cpp:281-285: n == 0 first, then NULL arguments
    assert init(ctypes.byref(h), 0, None, None, None) == S.ILLEGAL_ARGUMENT
    assert init(None, 1, a.ctypes.data, ja.ctypes.data, ia.ctypes.data) == S.NULL_ARGUMENT
    assert init(ctypes.byref(h), 1, None, ja.ctypes.data, ia.ctypes.data) == S.NULL_ARGUMENT
    assert init(ctypes.byref(h), 1, a.ctypes.data, None, ia.ctypes.data) == S.NULL_ARGUMENT
    assert init(ctypes.byref(h), 1, a.ctypes.data, ja.ctypes.data, None) == S.NULL_ARGUMENT
    # structural checks before the arrays reach the device
    bad_col = np.array([1], jt)  # column 1 in a 1x1 matrix
    assert init(ctypes.byref(h), 1, a.ctypes.data, bad_col.ctypes.data, ia.ctypes.data) == S.ILLEGAL_ARGUMENT
    bad_off = np.array([1, 0], it)  # decreasing row offsets
    assert init(ctypes.byref(h), 1, a.ctypes.data, ja.ctypes.data, bad_off.ctypes.data) == S.ILLEGAL_ARGUMENT
    if np.dtype(jt).kind == "i":
        neg = np.array([-1], jt)
        assert init(ctypes.byref(h), 1, a.ctypes.data, neg.ctypes.data, ia.ctypes.data) == S.ILLEGAL_ARGUMENT
    assert not h.value


@pytest.mark.parametrize("variant", ["sii", "dzz"])
def test_null_handle_accessors(S, variant):
    out = ctypes.c_double() if variant[0] == "d" else ctypes.c_float()
    sz = ctypes.c_size_t()
    for name, args in [("get_size", (ctypes.byref(sz),)), ("get_tolerance", (ctypes.byref(out),)),
                       ("get_residual", (ctypes.byref(out),)), ("get_iter_count", (ctypes.byref(sz),)),
                       ("set_tolerance", (1e-6,)), ("set_max_iter_count", (10,)),
                       ("set_preconditioner", (None, None)), ("get_preconditioner", (None, None)),
                       ("solve", (None, None)), ("destroy", ())]:
        assert S.fn(name, variant)(None, *args) == S.NULL_ARGUMENT, name
