"""ctypes binding of the transition CG C ABI (alp_hip_solver.h, the interface of
ALP's include/transition/solver.h) exported by libalp_hip.so.

    cg = SparseCG("dii", ia, ja, a)      # copies the CRS to the device
    cg.tolerance = 1e-10
    rc = cg.solve(x, b)                  # x: initial guess in, solution out
    cg.iterations, cg.residual

Variant names are ALP's: s/d value type, then the ja and ia integer types
(i = int32, z = size_t).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import lib as _lib

NO_ERROR, NULL_ARGUMENT, ILLEGAL_ARGUMENT, OUT_OF_MEMORY, FAILED, UNKNOWN = range(6)

_P = ctypes.c_void_p
_S = ctypes.c_size_t
VARIANTS = {
    # name: (value dtype, ja dtype, ia dtype)
    "sii": (np.float32, np.int32, np.int32),
    "siz": (np.float32, np.int32, np.uint64),
    "szz": (np.float32, np.uint64, np.uint64),
    "dii": (np.float64, np.int32, np.int32),
    "diz": (np.float64, np.int32, np.uint64),
    "dzz": (np.float64, np.uint64, np.uint64),
}
PRECOND_F64 = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                               ctypes.c_void_p)
PRECOND_F32 = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                               ctypes.c_void_p)


def fn(name: str, variant: str):
    """The raw C function sparse_cg_<name>_<variant> with its argument types set."""
    f = getattr(_lib(), f"sparse_cg_{name}_{variant}")
    real = ctypes.c_double if variant[0] == "d" else ctypes.c_float
    f.restype = ctypes.c_int
    f.argtypes = {
        "init": [_P, _S, _P, _P, _P],
        "get_size": [_P, _P],
        "get_tolerance": [_P, _P],
        "set_tolerance": [_P, real],
        "set_max_iter_count": [_P, _S],
        "get_residual": [_P, _P],
        "get_iter_count": [_P, _P],
        "get_preconditioner": [_P, _P, _P],
        "set_preconditioner": [_P, _P, _P],
        "solve": [_P, _P, _P],
        "destroy": [_P],
    }[name]
    return f


def _arr(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data


class SparseCG:
    """One sparse_cg handle of the given variant."""

    def __init__(self, variant: str, ia, ja, a):
        self.variant = variant
        self.dtype, jt, it = VARIANTS[variant]
        self.ia, p_ia = _arr(ia, it)
        self.ja, p_ja = _arr(ja, jt)
        self.a, p_a = _arr(a, self.dtype)
        self.n = len(self.ia) - 1
        self.h = ctypes.c_void_p()
        self._cb = None
        rc = fn("init", variant)(ctypes.byref(self.h), self.n, p_a, p_ja, p_ia)
        if rc != NO_ERROR:
            self.h = ctypes.c_void_p()
            raise RuntimeError(f"sparse_cg_init_{variant} returned {rc}")

    def _get(self, name, ctype):
        v = ctype()
        rc = fn(name, self.variant)(self.h, ctypes.byref(v))
        if rc != NO_ERROR:
            raise RuntimeError(f"sparse_cg_{name}_{self.variant} returned {rc}")
        return v.value

    @property
    def real(self):
        return ctypes.c_double if self.variant[0] == "d" else ctypes.c_float

    @property
    def tolerance(self):
        return self._get("get_tolerance", self.real)

    @tolerance.setter
    def tolerance(self, tol):
        assert fn("set_tolerance", self.variant)(self.h, tol) == NO_ERROR

    def set_max_iter(self, m: int):
        assert fn("set_max_iter_count", self.variant)(self.h, m) == NO_ERROR

    @property
    def size(self):
        return self._get("get_size", _S)

    @property
    def residual(self):
        return self._get("get_residual", self.real)

    @property
    def iterations(self):
        return self._get("get_iter_count", _S)

    def set_preconditioner(self, pyfn=None):
        """pyfn(z, r) writes z = M^-1 r into the numpy view z; returning an int
        other than 0 reports failure. None removes the preconditioner."""
        if pyfn is None:
            self._cb = None
            return fn("set_preconditioner", self.variant)(self.h, None, None)
        proto = PRECOND_F64 if self.variant[0] == "d" else PRECOND_F32
        n = self.n

        def cb(z, r, data):
            rv = pyfn(np.ctypeslib.as_array(z, (n,)), np.ctypeslib.as_array(r, (n,)))
            return rv if isinstance(rv, int) else 0

        self._cb = proto(cb)
        return fn("set_preconditioner", self.variant)(self.h, ctypes.cast(self._cb, _P), None)

    def solve(self, x: np.ndarray, b) -> int:
        assert x.dtype == self.dtype and x.flags.c_contiguous and len(x) == self.n
        b, pb = _arr(b, self.dtype)
        return fn("solve", self.variant)(self.h, x.ctypes.data, pb)

    def close(self):
        if self.h:
            fn("destroy", self.variant)(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
