"""TEST INFRASTRUCTURE ONLY -- ctypes bindings to the CPU oracle.

Two libraries are bound here:

* ``liboracle.so`` -- the project's plain-C restatement of ALP/GraphBLAS's
  sequential ``reference`` backend for the SpMV / CG / HPCG / PageRank / k-hop
  path (``oracle/alp_oracle.c``; every function cites the reference file:line it
  follows).
* ``_ref/libalpref_{seq,omp}.so`` -- ALP's own ``reference`` / ``reference_omp``
  backends compiled from the reference's sources behind a C shim
  (``oracle/ref_alp.cpp``), present only where those sources were available at
  build time.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the checker or the CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.c_void_p
_S = ctypes.c_size_t
_D = ctypes.c_double
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and oracle/_ref when the reference sources exist),
    every source on every call (make -B)."""
    out = subprocess.run(["make", "-B", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_stencil_nnz.restype = _S
        L.orc_stencil_nnz.argtypes = [_S, _S, _S]
        L.orc_stencil_crs.argtypes = [_S, _S, _S, _D, _D, _P, _P, _P]
        L.orc_coarsener_rows.restype = _S
        L.orc_coarsener_rows.argtypes = [_S, _S, _S]
        L.orc_coarsener_crs.argtypes = [_S, _S, _S, _P, _P, _P]
        L.orc_coo_to_compressed.argtypes = [_S, _S, _S, _P, _P, _P, ctypes.c_int, _P, _P, _P]
        L.orc_spmv_f64.argtypes = [_S, _P, _P, _P, _P, _P]
        L.orc_spmv_f32.argtypes = [_S, _P, _P, _P, _P, _P]
        L.orc_spmv_lorland.argtypes = [_S, _P, _P, _P, _P, _P, ctypes.c_int]
        L.orc_dot.restype = _D
        L.orc_dot.argtypes = [_S, _P, _P]
        L.orc_cg.argtypes = [_S, _P, _P, _P, _P, _P, _S, _D, _P, _P, _P]
        L.orc_hpcg_levels.argtypes = [_S, _S, _S, _S, _S, _P]
        L.orc_hpcg.argtypes = [_P, _P, _P, _S, _D, ctypes.c_int, _P, _P, _P]
        L.orc_mg_vcycle.argtypes = [_P, _P, _P]
        L.orc_pagerank.argtypes = [_S, _P, _P, _P, _P, _D, _D, _S, _P, _P, _P]
        L.orc_bfs_levels.argtypes = [_S, _P, _P, _U32, _P]
        L.orc_khop.argtypes = [_S, _P, _P, _U32, _S, _P]
        L.orc_rmat_edges.argtypes = [ctypes.c_uint, _S, _D, _D, _D, _U64, _P, _P]
        L.orc_pcg_f64.argtypes = [_S, _P, _P, _P, _P, _P, _S, _D, ctypes.c_void_p, _P, _P, _P]
        L.orc_pcg_f32.argtypes = [_S, _P, _P, _P, _P, _P, _S, ctypes.c_float, ctypes.c_void_p, _P, _P, _P]
        L.orc_mxm.argtypes = [_S, _S, _P, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_int, _P, _P, _P, _S, _P]
        L.orc_mix64.restype = _U64
        L.orc_mix64.argtypes = [_U64]
        _lib = L
    return _lib


# ---------------------------------------------------------------------------
# storage
# ---------------------------------------------------------------------------

class CSR:
    """Compressed storage in the reference's layout (col_start / row_index / values)."""

    def __init__(self, nrows, ncols, off, idx, val):
        self.nrows, self.ncols = int(nrows), int(ncols)
        self.off = np.ascontiguousarray(off, dtype=np.uint64)
        self.idx = np.ascontiguousarray(idx, dtype=np.uint32)
        self.val = None if val is None else np.ascontiguousarray(val, dtype=np.float64)

    @property
    def nnz(self) -> int:
        return int(self.off[-1])

    def dense(self) -> np.ndarray:
        A = np.zeros((self.nrows, self.ncols))
        for i in range(self.nrows):
            for k in range(int(self.off[i]), int(self.off[i + 1])):
                A[i, self.idx[k]] += 1.0 if self.val is None else self.val[k]
        return A


def stencil(nx: int, ny: int, nz: int, diag: float = 26.0, offdiag: float = -1.0) -> CSR:
    L = lib()
    n = nx * ny * nz
    nnz = L.orc_stencil_nnz(nx, ny, nz)
    off = np.zeros(n + 1, np.uint64)
    idx = np.zeros(nnz, np.uint32)
    val = np.zeros(nnz, np.float64)
    L.orc_stencil_crs(nx, ny, nz, diag, offdiag, _ptr(off), _ptr(idx), _ptr(val))
    return CSR(n, n, off, idx, val)


def coarsener(nxf: int, nyf: int, nzf: int) -> CSR:
    L = lib()
    nc = L.orc_coarsener_rows(nxf, nyf, nzf)
    off = np.zeros(nc + 1, np.uint64)
    idx = np.zeros(nc, np.uint32)
    val = np.zeros(nc, np.float64)
    L.orc_coarsener_crs(nxf, nyf, nzf, _ptr(off), _ptr(idx), _ptr(val))
    return CSR(nc, nxf * nyf * nzf, off, idx, val)


def coo_to_compressed(nrows, ncols, I, J, V=None, which=0) -> CSR:
    L = lib()
    I = np.ascontiguousarray(I, np.uint32)
    J = np.ascontiguousarray(J, np.uint32)
    nnz = len(I)
    m = nrows if which == 0 else ncols
    off = np.zeros(m + 1, np.uint64)
    idx = np.zeros(max(nnz, 1), np.uint32)
    val = np.zeros(max(nnz, 1), np.float64)
    Vp = None
    if V is not None:
        V = np.ascontiguousarray(V, np.float64)
        Vp = _ptr(V)
    rc = L.orc_coo_to_compressed(nrows, ncols, nnz, _ptr(I), _ptr(J), Vp, which,
                                 _ptr(off), _ptr(idx), _ptr(val))
    if rc != 0:
        raise ValueError("coordinate out of range")
    return CSR(nrows if which == 0 else ncols, ncols if which == 0 else nrows,
               off, idx[:nnz], val[:nnz] if V is not None else None)


# ---------------------------------------------------------------------------
# kernels and algorithms
# ---------------------------------------------------------------------------

def spmv(A: CSR, x: np.ndarray, y: np.ndarray | None = None) -> np.ndarray:
    """y <- y + A x (plus-times, f64 or f32 by x's dtype), reference fold order."""
    L = lib()
    if x.dtype == np.float32:
        y = np.zeros(A.nrows, np.float32) if y is None else np.array(y, np.float32)
        v = np.ascontiguousarray(A.val, np.float32)
        L.orc_spmv_f32(A.nrows, _ptr(A.off), _ptr(A.idx), _ptr(v), _ptr(np.ascontiguousarray(x)), _ptr(y))
        return y
    y = np.zeros(A.nrows) if y is None else np.array(y, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    L.orc_spmv_f64(A.nrows, _ptr(A.off), _ptr(A.idx), _ptr(A.val), _ptr(x), _ptr(y))
    return y


def spmv_lorland(A: CSR, x: np.ndarray, y: np.ndarray, mask=None, invert_mask=False):
    L = lib()
    x = np.ascontiguousarray(x, np.uint8)
    y = np.array(y, np.uint8)
    mp = None
    if mask is not None:
        mask = np.ascontiguousarray(mask, np.uint8)
        mp = _ptr(mask)
    L.orc_spmv_lorland(A.nrows, _ptr(A.off), _ptr(A.idx), _ptr(x), _ptr(y), mp, int(invert_mask))
    return y


def dot(x: np.ndarray, y: np.ndarray) -> float:
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return float(lib().orc_dot(len(x), _ptr(x), _ptr(y)))


def cg(A: CSR, b: np.ndarray, x0: np.ndarray, max_iter: int, tol: float):
    """Returns (rc, x, iterations, residual=||r||^2, history of ||r||^2)."""
    L = lib()
    x = np.array(x0, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    it = _S()
    res = _D()
    hist = np.zeros(max_iter)
    rc = L.orc_cg(A.nrows, _ptr(A.off), _ptr(A.idx), _ptr(A.val), _ptr(b), _ptr(x),
                  max_iter, tol, ctypes.byref(it), ctypes.byref(res), _ptr(hist))
    return rc, x, it.value, res.value, hist[: it.value]


class HpcgCfg(ctypes.Structure):
    _fields_ = [("levels", _S), ("nx", _S * 8), ("ny", _S * 8), ("nz", _S * 8),
                ("colors", _S), ("geometric", ctypes.c_int), ("presmooth", _S),
                ("postsmooth", _S)]


def hpcg_cfg(nx, ny, nz, max_levels, colors=8, geometric=False, pre=1, post=1, min_phys=4):
    cfg = HpcgCfg()
    lib().orc_hpcg_levels(nx, ny, nz, max_levels, min_phys, ctypes.byref(cfg))
    cfg.colors = colors
    cfg.geometric = int(geometric)
    cfg.presmooth = pre
    cfg.postsmooth = post
    return cfg


def hpcg(cfg: HpcgCfg, b: np.ndarray, x0: np.ndarray, max_iter: int, tol: float,
         with_precond: bool = True):
    """Returns (x, iterations, ||r||, history of ||r||)."""
    L = lib()
    x = np.array(x0, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    it = _S()
    res = _D()
    hist = np.zeros(max_iter)
    rc = L.orc_hpcg(ctypes.byref(cfg), _ptr(b), _ptr(x), max_iter, tol, int(with_precond),
                    ctypes.byref(it), ctypes.byref(res), _ptr(hist))
    if rc != 0:
        raise RuntimeError("orc_hpcg failed")
    return x, it.value, res.value, hist[: it.value]


def mg_vcycle(cfg: HpcgCfg, r: np.ndarray) -> np.ndarray:
    r = np.ascontiguousarray(r, np.float64)
    z = np.zeros_like(r)
    if lib().orc_mg_vcycle(ctypes.byref(cfg), _ptr(r), _ptr(z)) != 0:
        raise RuntimeError("orc_mg_vcycle failed")
    return z


def pagerank(crs: CSR, ccs: CSR, alpha=0.85, conv=1e-7, max_iter=1000):
    L = lib()
    n = crs.nrows
    pr = np.zeros(n)
    it = _S()
    q = _D()
    rc = L.orc_pagerank(n, _ptr(crs.off), _ptr(crs.idx), _ptr(ccs.off), _ptr(ccs.idx), alpha,
                        conv, max_iter, _ptr(pr), ctypes.byref(it), ctypes.byref(q))
    return rc, pr, it.value, q.value


def mxm(A: CSR, B: CSR, ring: int = 0, C: CSR | None = None) -> CSR:
    """C (+)= A (x) B in mxm_generic's order (reference/blas3.hpp:1148-1725): ring 0
    plus-times, 1 min-plus, 2 lor-land (pattern result). A None value array is a
    pattern operand (entries = the multiplicative identity). Pass transposes as
    their CCS (the rows of A^T)."""
    assert A.ncols == B.nrows
    m, n = A.nrows, B.ncols
    ca = C is not None and C.nnz > 0
    # products per row bound the output
    bl = np.diff(B.off.astype(np.int64))
    cs = np.concatenate([[0], np.cumsum(bl[A.idx[:A.nnz].astype(np.int64)])])
    cap = int(cs[-1]) + (C.nnz if ca else 0) + 1
    off = np.zeros(m + 1, np.uint64)
    idx = np.zeros(cap, np.uint32)
    val = np.zeros(cap, np.float64)
    nz = _S()
    L = lib()
    rc = L.orc_mxm(m, n, _ptr(A.off), _ptr(A.idx), None if A.val is None else _ptr(A.val),
                   _ptr(B.off), _ptr(B.idx), None if B.val is None else _ptr(B.val),
                   _ptr(C.off) if ca else None, _ptr(C.idx) if ca else None,
                   _ptr(C.val) if ca and C.val is not None else None, ring,
                   _ptr(off), _ptr(idx), _ptr(val), cap, ctypes.byref(nz))
    assert rc == 0, rc
    k = nz.value
    return CSR(m, n, off, idx[:k], None if ring == 2 else val[:k])


def bfs_levels(A: CSR, source: int) -> np.ndarray:
    lev = np.zeros(A.nrows, np.int32)
    lib().orc_bfs_levels(A.nrows, _ptr(A.off), _ptr(A.idx), source, _ptr(lev))
    return lev


def khop(A: CSR, source: int, k: int) -> np.ndarray:
    out = np.zeros(A.nrows, np.uint8)
    lib().orc_khop(A.nrows, _ptr(A.off), _ptr(A.idx), source, k, _ptr(out))
    return out


def rmat_edges(scale: int, nedges: int, a=0.57, b=0.19, c=0.19, seed=1):
    src = np.zeros(nedges, np.uint32)
    dst = np.zeros(nedges, np.uint32)
    lib().orc_rmat_edges(scale, nedges, a, b, c, seed, _ptr(src), _ptr(dst))
    return src, dst


def mix64(x: int) -> int:
    return int(lib().orc_mix64(x))


def symmetrize_dedup(n: int, src: np.ndarray, dst: np.ndarray):
    """Symmetrise, drop self-loops, deduplicate; returns (I, J) sorted by (I, J)."""
    s = np.concatenate([src, dst]).astype(np.uint64)
    d = np.concatenate([dst, src]).astype(np.uint64)
    keep = s != d
    key = np.unique(s[keep] * np.uint64(n) + d[keep])
    return (key // np.uint64(n)).astype(np.uint32), (key % np.uint64(n)).astype(np.uint32)


PRECOND_F64 = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)
PRECOND_F32 = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                               ctypes.POINTER(ctypes.c_float), ctypes.c_void_p)


def precond_callback(fn, n: int, dtype=np.float64):
    """Wraps fn(z, r) (numpy views of length n; writes z) as a C preconditioner;
    an int returned by fn is the callback's status, anything else means 0."""
    proto = PRECOND_F64 if np.dtype(dtype) == np.float64 else PRECOND_F32

    def cb(z, r, data):
        rv = fn(np.ctypeslib.as_array(z, (n,)), np.ctypeslib.as_array(r, (n,)))
        return rv if isinstance(rv, int) else 0

    return proto(cb)


def pcg(off, idx, val, b, x0, max_iter: int, tol: float, precond=None):
    """The transition solver's CG (preconditioned when precond is a callback
    from precond_callback) in val's precision on a caller's CRS; returns
    (status, x, iterations, residual), status 0 converged, 1 iteration limit,
    2 preconditioner failure."""
    dt = np.dtype(val.dtype)
    off = np.ascontiguousarray(off, np.uint64)
    idx = np.ascontiguousarray(idx, np.uint32)
    val = np.ascontiguousarray(val, dt)
    b = np.ascontiguousarray(b, dt)
    x = np.array(x0, dtype=dt)
    it = _S()
    if dt == np.float64:
        res = _D()
        f = lib().orc_pcg_f64
    else:
        res = ctypes.c_float()
        f = lib().orc_pcg_f32
    rc = f(len(off) - 1, _ptr(off), _ptr(idx), _ptr(val), _ptr(b), _ptr(x), max_iter, tol,
           ctypes.cast(precond, ctypes.c_void_p) if precond is not None else None, None,
           ctypes.byref(it), ctypes.byref(res))
    if rc < 0:
        raise MemoryError("orc_pcg")
    return rc, x, it.value, res.value


# ---------------------------------------------------------------------------
# ALP's own reference backend (oracle/_ref), when it was built
# ---------------------------------------------------------------------------

_refs: dict = {}


def ref_available(flavour: str = "seq") -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", f"libalpref_{flavour}.so"))


def ref_solver() -> ctypes.CDLL:
    """ALP's transition CG C ABI on the reference backend (oracle/_ref)."""
    if "solver" not in _refs:
        path = os.path.join(HERE, "_ref", "libalpref_solver.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        _refs["solver"] = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    return _refs["solver"]


def ref(flavour: str = "seq") -> ctypes.CDLL:
    if flavour not in _refs:
        path = os.path.join(HERE, "_ref", f"libalpref_{flavour}.so")
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        R = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        R.ref_stencil_nnz.argtypes = [_S, _S, _S, _P]
        R.ref_stencil_storage.argtypes = [_S, _S, _S, ctypes.c_int, _P, _P, _P]
        R.ref_stencil_mxv.argtypes = [_S, _S, _S, _P, _P]
        R.ref_stencil_mxv_transposed.argtypes = [_S, _S, _S, _P, _P]
        R.ref_dot.argtypes = [_S, _P, _P, _P]
        R.ref_stencil_cg.argtypes = [_S, _S, _S, _P, _P, _S, _D, _P, _P, _P]
        R.ref_hpcg.argtypes = [_S, _S, _S, _S, _S, _S, _S, _S, _D, ctypes.c_int, _P, _P, _P]
        R.ref_hpcg_levels.argtypes = [_S, _S, _S, _S, _P]
        R.ref_pagerank.argtypes = [_S, _S, _P, _P, _D, _D, _S, _P, _P, _P]
        R.ref_knn.argtypes = [_S, _S, _P, _P, _S, _S, _P]
        R.ref_mxm.argtypes = [_S, _S, _S, _S, _P, _P, _P, _S, _P, _P, _P, _S, _P, _P, _P, ctypes.c_int,
                              ctypes.c_int, _P, _P, _P, _P, _P, _P, _S, _P]
        R.ref_time_stencil_mxv.restype = _D
        R.ref_time_stencil_mxv.argtypes = [_S, _S, _S, _S]
        R.ref_time_stencil_cg.restype = _D
        R.ref_time_stencil_cg.argtypes = [_S, _S, _S, _S, _D, _P]
        R.ref_time_hpcg.restype = _D
        R.ref_time_hpcg.argtypes = [_S, _S, _S, _S, _S, _S, _P]
        _refs[flavour] = R
    return _refs[flavour]
