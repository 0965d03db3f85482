"""alp_amd -- Python binding of the MI355X-native `hip` backend for ALP/GraphBLAS.

The product is ``libalp_hip.so`` (hand-written gfx950 kernels behind the C ABI
declared in ``include/alp_hip.h``). This module binds that ABI with ctypes for
tests, benchmarks and Python callers. It never falls back to a CPU path: if the
library is missing, importing :func:`lib` raises.

Objects mirror ALP's containers (include/graphblas/base/{vector,matrix}.hpp):
:class:`Vector` owns a device array, :class:`Matrix` owns device CRS/CCS
storage; the functions named after grb:: primitives (``mxv``, ``vxm``, ``dot``,
...) follow the reference's argument order and return codes (grb::RC values).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libalp_hip.so")

# grb::RC (include/graphblas/rc.hpp:47-155)
SUCCESS, PANIC, OUTOFMEM, MISMATCH, OVERLAP, OVERFLW, UNSUPPORTED, ILLEGAL, FAILED = range(9)
RC_NAMES = ["SUCCESS", "PANIC", "OUTOFMEM", "MISMATCH", "OVERLAP", "OVERFLW", "UNSUPPORTED",
            "ILLEGAL", "FAILED"]

# semirings (alp_semiring_t)
PLUS_TIMES_F64, PLUS_TIMES_F32, LOR_LAND_BOOL, PLUS_PATTERN_F64, MIN_PLUS_F64 = range(5)

# descriptors (include/graphblas/descriptors.hpp:63-267)
INVERT_MASK = 1
TRANSPOSE_MATRIX = 2
STRUCTURAL = 8
DENSE = 16
ADD_IDENTITY = 32
SAFE_OVERLAP = 1024
TRANSPOSE_LEFT = 2048
TRANSPOSE_RIGHT = 4096
OVERWRITE = 1 << 20

VAL_F64, VAL_F32, VAL_PATTERN = 0, 1, 2
MTX_PATTERN, MTX_SYMMETRIC = 1, 2
MAT_WITH_CCS, MAT_WITH_F32 = 1, 2
GRAPH_SYMMETRIZE, GRAPH_DROP_SELF_LOOPS = 1, 2


class AlpError(RuntimeError):
    def __init__(self, rc: int, what: str):
        self.rc = rc
        super().__init__(f"{what} returned {RC_NAMES[rc] if 0 <= rc < 9 else rc}")


def build(quiet: bool = True) -> None:
    """Compile libalp_hip.so and the grb:: test programs for gfx950 in-tree (hipcc
    cross-compiles without a GPU). Every source is compiled on every call (make
    -B), so the call checks that the tree builds, whatever is already built."""
    out = subprocess.run(["make", "-B", "-C", HERE, f"-j{min(16, os.cpu_count() or 4)}"],
                         capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("libalp_hip build failed:\n" + out.stdout[-4000:] + out.stderr[-8000:])
    if not quiet:
        print(out.stdout)


_P = ctypes.c_void_p
_S = ctypes.c_size_t
_D = ctypes.c_double
_I = ctypes.c_int
_U = ctypes.c_uint
_RC = ctypes.c_int

# every exported entry point: name -> (restype, argtypes)
SIGNATURES = {
    "alp_hip_init": (_RC, [_I]),
    "alp_hip_finalize": (_RC, []),
    "alp_hip_set_stream": (_RC, [_P]),
    "alp_hip_sync": (_RC, []),
    "alp_hip_version": (ctypes.c_char_p, []),
    "alp_hip_device_count": (_I, []),
    "alp_hip_timer_start": (_RC, []),
    "alp_hip_timer_stop": (_RC, [ctypes.POINTER(_D)]),
    "alp_hip_malloc": (_RC, [ctypes.POINTER(_P), _S]),
    "alp_hip_free": (_RC, [_P]),
    "alp_hip_memcpy_h2d": (_RC, [_P, _P, _S]),
    "alp_hip_memcpy_d2h": (_RC, [_P, _P, _S]),
    "alp_hip_memcpy_d2d": (_RC, [_P, _P, _S]),
    "alp_hip_matrix_create_crs": (_RC, [ctypes.POINTER(_P), _S, _S, _S, _P, _P, _P, _I, _U]),
    "alp_hip_matrix_create_coo": (_RC, [ctypes.POINTER(_P), _S, _S, _S, _P, _P, _P, _I, _U]),
    "alp_hip_matrix_stencil27": (_RC, [ctypes.POINTER(_P), _S, _S, _S, _D, _D, _U]),
    "alp_hip_matrix_destroy": (_RC, [_P]),
    "alp_hip_matrix_info": (_RC, [_P, ctypes.POINTER(_S), ctypes.POINTER(_S), ctypes.POINTER(_S)]),
    "alp_hip_matrix_download": (_RC, [_P, _I, _P, _P, _P]),
    "alp_hip_matrix_set_kernel": (_RC, [_P, _I, _I, _I]),
    "alp_hip_matrix_get_kernel": (_RC, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "alp_hip_matrix_autotune": (_RC, [_P]),
    "alp_hip_mtx_read": (_RC, [ctypes.c_char_p, ctypes.POINTER(_S), ctypes.POINTER(_S), ctypes.POINTER(_S),
                               ctypes.POINTER(_S), _P, _P, _P, ctypes.POINTER(_I)]),
    "alp_hip_matrix_read_mtx": (_RC, [ctypes.POINTER(_P), ctypes.c_char_p, _I, _U]),
    "alp_hip_matrix_layout": (_RC, [_P, _I, ctypes.POINTER(_S), ctypes.POINTER(_S), ctypes.POINTER(_S)]),
    "alp_hip_get_stream": (_RC, [ctypes.POINTER(_P)]),
    "alp_hip_matrix_storage": (_RC, [_P, _I, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "alp_hip_matrix_device_bytes": (_RC, [_P, ctypes.POINTER(_S)]),
    "alp_hip_matrix_create_coo_map": (_RC, [ctypes.POINTER(_P), _S, _S, _S, _P, _P, _U, _P, _P]),
    "alp_hip_matrix_create_device": (_RC, [ctypes.POINTER(_P), _S, _S, _S, _P, _P, _P, _I, _U, _P]),
    "alp_hip_mxm": (_RC, [ctypes.POINTER(_P), _P, _P, _I, _U]),
    "alp_hip_mxv": (_RC, [_P, _I, _U, _P, _P, _P]),
    "alp_hip_vxm": (_RC, [_P, _I, _U, _P, _P, _P]),
    "alp_hip_dot_f64": (_RC, [_S, _P, _P, ctypes.POINTER(_D)]),
    "alp_hip_dot_f32": (_RC, [_S, _P, _P, ctypes.POINTER(_D)]),
    "alp_hip_set_f64": (_RC, [_S, _P, _D]),
    "alp_hip_set_f32": (_RC, [_S, _P, ctypes.c_float]),
    "alp_hip_set_u8": (_RC, [_S, _P, ctypes.c_uint8]),
    "alp_hip_axpy_f64": (_RC, [_S, _D, _P, _P]),
    "alp_hip_xpay_f64": (_RC, [_S, _D, _P, _P, _P]),
    "alp_hip_ewise_f64": (_RC, [_S, _I, _P, _P, _P]),
    "alp_hip_scale_f64": (_RC, [_S, _D, _P]),
    "alp_hip_sum_f64": (_RC, [_S, _P, ctypes.POINTER(_D)]),
    "alp_hip_convert_f64_to_f32": (_RC, [_S, _P, _P]),
    "alp_hip_convert_f32_to_f64": (_RC, [_S, _P, _P]),
    "alp_hip_cg": (_RC, [_P, _P, _P, _S, _D, _I, ctypes.POINTER(_S), ctypes.POINTER(_D)]),
    "alp_hip_cg_transposed": (_RC, [_P, _P, _P, _S, _D, _I, ctypes.POINTER(_S), ctypes.POINTER(_D)]),
    "alp_hip_cg_noalloc": (_RC, [_P, _P, _P, _S, _D, ctypes.POINTER(_S), ctypes.POINTER(_D)]),
    "alp_hip_matrix_conductance27": (_RC, [ctypes.POINTER(_P), _S, _S, _S, _P, _D, _U]),
    "alp_hip_grad_conductance27": (_RC, [_S, _S, _S, _P, _P, _P]),
    "alp_hip_grad_matrix": (_RC, [_P, _P, _P, _P]),
    "alp_hip_rmat_edges": (_RC, [_U, _S, _D, _D, _D, ctypes.c_uint64, _P, _P]),
    "alp_hip_graph_from_edges": (_RC, [ctypes.POINTER(_P), _S, _S, _P, _P, _U, ctypes.POINTER(_S)]),
    "alp_hip_max_degree_vertex": (_RC, [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(_S)]),
    "alp_hip_bfs": (_RC, [_P, ctypes.c_uint32, _P, ctypes.POINTER(_S), ctypes.POINTER(_S)]),
    "alp_hip_pagerank": (_RC, [_P, _P, _D, _D, _S, ctypes.POINTER(_S), ctypes.POINTER(_D)]),
    "alp_hip_khop": (_RC, [_P, ctypes.c_uint32, _S, _P]),
    "alp_hip_batch_create_laplace7": (_RC, [ctypes.POINTER(_P), _S, _P, _P, _P, _P, _P, _P, _S]),
    "alp_hip_batch_create_crs": (_RC, [ctypes.POINTER(_P), _S, _P, _P, _P, _P, _S]),
    "alp_hip_batch_destroy": (_RC, [_P]),
    "alp_hip_batch_info": (_RC, [_P, ctypes.POINTER(_S), ctypes.POINTER(_S), _P]),
    "alp_hip_batch_layout": (_RC, [_P, ctypes.POINTER(_S), ctypes.POINTER(_S)]),
    "alp_hip_batch_spmm": (_RC, [_P, _P, _P]),
    "alp_hip_batch_autotune": (_RC, [_P, ctypes.POINTER(_I)]),
    "alp_hip_batch_solve": (_RC, [_P, _P, _P, _D, _S, _I, _P, _P, ctypes.POINTER(_S)]),
    "alp_hip_hpcg_create": (_RC, [ctypes.POINTER(_P), _S, _S, _S, _S, _S, _S, _S, _U]),
    "alp_hip_hpcg_destroy": (_RC, [_P]),
    "alp_hip_hpcg_info": (_RC, [_P, ctypes.POINTER(_S), ctypes.POINTER(_S), ctypes.POINTER(_S),
                                 ctypes.POINTER(_I)]),
    "alp_hip_hpcg_vcycle": (_RC, [_P, _P, _P]),
    "alp_hip_hpcg_solve": (_RC, [_P, _P, _P, _S, _D, _I, _I, ctypes.POINTER(_S), ctypes.POINTER(_D),
                                  ctypes.POINTER(_S)]),
}

_lib = None


def hip_runtimes() -> list:
    """Paths of the HIP runtime libraries (libamdhip64) mapped into this process."""
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                p = line.split()[-1]
                if "libamdhip64" in os.path.basename(p):
                    paths.add(os.path.realpath(p))
    except OSError:
        pass
    return sorted(paths)


def lib() -> ctypes.CDLL:
    """The loaded libalp_hip.so (raises if it was not built: there is no fallback).

    libalp_hip.so needs the HIP runtime by its soname (libamdhip64.so.7); when
    PyTorch (ROCm) is imported first, the library binds to the runtime torch
    loaded, so one process holds one runtime. Loading this library first and
    torch afterwards would load a second runtime, which cannot use the device:
    import torch (or alp_amd.autograd) before the first library call when both
    are used."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `make -C alp_amd` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str) -> None:
    if rc != SUCCESS:
        raise AlpError(rc, what)


def init(device: int = 0) -> None:
    check(lib().alp_hip_init(device), "alp_hip_init")


def sync() -> None:
    check(lib().alp_hip_sync(), "alp_hip_sync")


def device_count() -> int:
    return int(lib().alp_hip_device_count())


# ---------------------------------------------------------------------------
# containers
# ---------------------------------------------------------------------------

class Vector:
    """A dense device vector (grb::Vector<D, hip>), owning hipMalloc'd memory."""

    def __init__(self, n: int, dtype=np.float64, fill=None):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = _P()
        check(lib().alp_hip_malloc(ctypes.byref(p), max(1, self.n * self.dtype.itemsize)), "alp_hip_malloc")
        self.ptr = p.value
        if fill is not None:
            self.set(fill)

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "Vector":
        a = np.ascontiguousarray(a)
        v = cls(a.size, a.dtype)
        v.upload(a)
        return v

    @property
    def nbytes(self) -> int:
        return self.n * self.dtype.itemsize

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(lib().alp_hip_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes), "memcpy_h2d")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        check(lib().alp_hip_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes), "memcpy_d2h")
        return out

    def set(self, v) -> None:
        L = lib()
        if self.dtype == np.float64:
            check(L.alp_hip_set_f64(self.n, self.ptr, float(v)), "set")
        elif self.dtype == np.float32:
            check(L.alp_hip_set_f32(self.n, self.ptr, float(v)), "set")
        elif self.dtype.itemsize == 1:
            check(L.alp_hip_set_u8(self.n, self.ptr, int(v)), "set")
        else:
            self.upload(np.full(self.n, v, self.dtype))

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib().alp_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Matrix:
    """Device-resident CRS (+ optional CCS) storage: grb::Matrix<D, hip>."""

    def __init__(self, handle: int):
        self.h = handle
        r, c, z = _S(), _S(), _S()
        check(lib().alp_hip_matrix_info(self.h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(z)),
              "matrix_info")
        self.nrows, self.ncols, self.nnz = r.value, c.value, z.value

    @classmethod
    def stencil27(cls, nx, ny, nz, diag=26.0, offdiag=-1.0, flags=0) -> "Matrix":
        h = _P()
        check(lib().alp_hip_matrix_stencil27(ctypes.byref(h), nx, ny, nz, diag, offdiag, flags),
              "matrix_stencil27")
        return cls(h.value)

    @classmethod
    def conductance27(cls, nx, ny, nz, theta, shift: float = 1e-2, flags: int = 0) -> "Matrix":
        """A(theta) of the differentiable solve; theta: a Vector or a device
        pointer to 13 nx ny nz doubles (read during the call only)."""
        h = _P()
        tp = theta.ptr if isinstance(theta, Vector) else theta
        check(lib().alp_hip_matrix_conductance27(ctypes.byref(h), nx, ny, nz, tp, shift, flags),
              "matrix_conductance27")
        return cls(h.value)

    @classmethod
    def from_crs(cls, nrows, ncols, off, idx, val=None, value_type=None, flags=0) -> "Matrix":
        off = np.ascontiguousarray(off, np.uint64)
        idx = np.ascontiguousarray(idx, np.uint32)
        if value_type is None:
            value_type = VAL_PATTERN if val is None else (VAL_F32 if np.asarray(val).dtype == np.float32 else VAL_F64)
        vp = None
        if val is not None:
            val = np.ascontiguousarray(val, np.float32 if value_type == VAL_F32 else np.float64)
            vp = val.ctypes.data
        h = _P()
        check(lib().alp_hip_matrix_create_crs(ctypes.byref(h), nrows, ncols, int(off[-1]),
                                              off.ctypes.data, idx.ctypes.data, vp, value_type, flags),
              "matrix_create_crs")
        return cls(h.value)

    @classmethod
    def from_coo(cls, nrows, ncols, I, J, V=None, value_type=None, flags=0) -> "Matrix":
        I = np.ascontiguousarray(I, np.uint32)
        J = np.ascontiguousarray(J, np.uint32)
        if value_type is None:
            value_type = VAL_PATTERN if V is None else (VAL_F32 if np.asarray(V).dtype == np.float32 else VAL_F64)
        vp = None
        if V is not None:
            V = np.ascontiguousarray(V, np.float32 if value_type == VAL_F32 else np.float64)
            vp = V.ctypes.data
        h = _P()
        check(lib().alp_hip_matrix_create_coo(ctypes.byref(h), nrows, ncols, len(I), I.ctypes.data,
                                              J.ctypes.data, vp, value_type, flags),
              "matrix_create_coo")
        return cls(h.value)

    @classmethod
    def from_mtx(cls, path: str, value_type=VAL_F64, flags=0) -> "Matrix":
        """A MatrixMarket coordinate file, built on the device (alp_hip_matrix_read_mtx)."""
        h = _P()
        check(lib().alp_hip_matrix_read_mtx(ctypes.byref(h), os.fsencode(path), value_type, flags),
              "matrix_read_mtx")
        return cls(h.value)

    def download(self, which=0, with_values=True, value_dtype=np.float64):
        nmaj = self.nrows if which == 0 else self.ncols
        off = np.zeros(nmaj + 1, np.uint64)
        idx = np.zeros(max(self.nnz, 1), np.uint32)
        val = np.zeros(max(self.nnz, 1), value_dtype) if with_values else None
        check(lib().alp_hip_matrix_download(self.h, which, off.ctypes.data, idx.ctypes.data,
                                            val.ctypes.data if with_values else None), "matrix_download")
        return off, idx[: self.nnz], (val[: self.nnz] if with_values else None)

    def set_kernel(self, kernel: int, lanes: int = 0, rows: int = 0) -> None:
        check(lib().alp_hip_matrix_set_kernel(self.h, kernel, lanes, rows), "matrix_set_kernel")

    def get_kernel(self):
        """(kernel, lanes, rows): 1 = CSR-vector with `lanes` per row and `rows` per group, 2 = CSR-stream,
        3 = CSR-wave, 4 = sliced ELL, 5 = sliced ELL with index-free diagonal slices."""
        k, l, r = _I(), _I(), _I()
        check(lib().alp_hip_matrix_get_kernel(self.h, ctypes.byref(k), ctypes.byref(l), ctypes.byref(r)),
              "matrix_get_kernel")
        return k.value, l.value, r.value

    def autotune(self):
        check(lib().alp_hip_matrix_autotune(self.h), "matrix_autotune")
        return self.get_kernel()

    def layout(self, which: int = 0):
        """(slices, padded entries, slices in diagonal form) of the CRS (0) or CCS (1) sliced layout."""
        a, b, c = _S(), _S(), _S()
        check(lib().alp_hip_matrix_layout(self.h, which, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)),
              "matrix_layout")
        return a.value, b.value, c.value

    def device_bytes(self) -> int:
        """Device memory the matrix holds (CRS, CCS and every derived layout)."""
        b = _S()
        check(lib().alp_hip_matrix_device_bytes(self.h, ctypes.byref(b)), "matrix_device_bytes")
        return b.value

    def free(self) -> None:
        if getattr(self, "h", None):
            lib().alp_hip_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# primitives (grb:: names)
# ---------------------------------------------------------------------------

def mxv(y: Vector, A: Matrix, x: Vector, ring: int = PLUS_TIMES_F64, descr: int = 0,
        mask: Vector | None = None) -> int:
    """y (+)= A (x) x (grb::mxv, reference/blas2.hpp:1946); returns the grb::RC code."""
    return lib().alp_hip_mxv(A.h, ring, descr, y.ptr, x.ptr, mask.ptr if mask is not None else None)


def vxm(y: Vector, x: Vector, A: Matrix, ring: int = PLUS_TIMES_F64, descr: int = 0,
        mask: Vector | None = None) -> int:
    """y (+)= x (x) A (grb::vxm, reference/blas2.hpp:1540-1694); returns the grb::RC code."""
    return lib().alp_hip_vxm(A.h, ring, descr, y.ptr, x.ptr, mask.ptr if mask is not None else None)


def mxm(A: Matrix, B: Matrix, ring: int = PLUS_TIMES_F64, descr: int = 0, C: Matrix | None = None) -> Matrix:
    """C (+)= op(A) (x) op(B) (grb::mxm, reference/blas3.hpp:1148-1725): a new matrix,
    or, when C is given and holds entries, C updated in place (C's handle is
    replaced by the result)."""
    h = _P(C.h if C is not None else None)
    check(lib().alp_hip_mxm(ctypes.byref(h), A.h, B.h, ring, descr), "mxm")
    if C is not None:
        C.h = h.value
        r, c, z = _S(), _S(), _S()
        check(lib().alp_hip_matrix_info(C.h, ctypes.byref(r), ctypes.byref(c), ctypes.byref(z)), "matrix_info")
        C.nrows, C.ncols, C.nnz = r.value, c.value, z.value
        return C
    return Matrix(h.value)


def dot(x: Vector, y: Vector, z: float = 0.0) -> float:
    """z + dot(x, y) with a deterministic reduction (grb::dot, reference/blas1.hpp:9520)."""
    out = _D(z)
    f = lib().alp_hip_dot_f32 if x.dtype == np.float32 else lib().alp_hip_dot_f64
    check(f(x.n, x.ptr, y.ptr, ctypes.byref(out)), "dot")
    return out.value


def mtx_read(path: str):
    """Host parse of a MatrixMarket coordinate file (alp_hip_mtx_read): (nrows, ncols,
    I, J, V, is_pattern), entries in the reference iterator's order (a symmetric
    file's off-diagonal entry followed by its mirror), 0-based indices."""
    r, c, z, e, kind = _S(), _S(), _S(), _S(), _I()
    p = os.fsencode(path)
    check(lib().alp_hip_mtx_read(p, ctypes.byref(r), ctypes.byref(c), ctypes.byref(z), ctypes.byref(e), None, None,
                                 None, ctypes.byref(kind)), "mtx_read")
    nnz = z.value
    I = np.zeros(max(nnz, 1), np.uint32)
    J = np.zeros(max(nnz, 1), np.uint32)
    V = np.zeros(max(nnz, 1), np.float64)
    check(lib().alp_hip_mtx_read(p, ctypes.byref(r), ctypes.byref(c), ctypes.byref(z), ctypes.byref(e),
                                 I.ctypes.data, J.ctypes.data, V.ctypes.data, ctypes.byref(kind)), "mtx_read")
    return r.value, c.value, I[:nnz], J[:nnz], V[:nnz], bool(kind.value & MTX_PATTERN)


def cg(A: Matrix, b, x, max_iter: int, tol: float, precision: int = 0, transpose: bool = False):
    """grb::algorithms::conjugate_gradient on the device (on A^T with transpose);
    b and x are Vectors or raw device pointers; returns (rc, iterations, ||r||^2)."""
    it, res = _S(), _D()
    f = lib().alp_hip_cg_transposed if transpose else lib().alp_hip_cg
    bp = b.ptr if isinstance(b, Vector) else b
    xp = x.ptr if isinstance(x, Vector) else x
    rc = f(A.h, bp, xp, max_iter, tol, precision, ctypes.byref(it), ctypes.byref(res))
    return rc, it.value, res.value


class Hpcg:
    """The HPCG multigrid hierarchy on the device (build_hpcg_system + hpcg)."""

    def __init__(self, nx, ny, nz, max_levels=3, colours=0, presmooth=1, postsmooth=1, f32=True):
        h = _P()
        check(lib().alp_hip_hpcg_create(ctypes.byref(h), nx, ny, nz, max_levels, colours, presmooth,
                                        postsmooth, MAT_WITH_F32 if f32 else 0), "hpcg_create")
        self.h = h.value
        n, nnz, lv, pr = _S(), _S(), _S(), _I()
        check(lib().alp_hip_hpcg_info(self.h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(lv),
                                      ctypes.byref(pr)), "hpcg_info")
        self.n, self.nnz, self.levels, self.proper = n.value, nnz.value, lv.value, bool(pr.value)

    def vcycle(self, r: Vector, z: Vector) -> None:
        check(lib().alp_hip_hpcg_vcycle(self.h, r.ptr, z.ptr), "hpcg_vcycle")

    def solve(self, b: Vector, x: Vector, max_iter: int, tol: float, precision: int = 0,
              with_precond: bool = True):
        """Returns (iterations, ||r||, refinement steps)."""
        it, res, outer = _S(), _D(), _S()
        check(lib().alp_hip_hpcg_solve(self.h, b.ptr, x.ptr, max_iter, tol, precision, int(with_precond),
                                       ctypes.byref(it), ctypes.byref(res), ctypes.byref(outer)), "hpcg_solve")
        return it.value, res.value, outer.value

    def free(self) -> None:
        if getattr(self, "h", None):
            lib().alp_hip_hpcg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# graphs
# ---------------------------------------------------------------------------

def rmat_edges(scale: int, nedges: int, a=0.57, b=0.19, c=0.19, seed=1):
    """Device R-MAT edge lists (src, dst) as uint32 Vectors."""
    src = Vector(nedges, np.uint32)
    dst = Vector(nedges, np.uint32)
    check(lib().alp_hip_rmat_edges(scale, nedges, a, b, c, seed, src.ptr, dst.ptr), "rmat_edges")
    return src, dst


def graph_from_edges(n: int, src: Vector, dst: Vector, flags=GRAPH_SYMMETRIZE | GRAPH_DROP_SELF_LOOPS) -> Matrix:
    h = _P()
    nnz = _S()
    check(lib().alp_hip_graph_from_edges(ctypes.byref(h), n, src.n, src.ptr, dst.ptr, flags, ctypes.byref(nnz)),
          "graph_from_edges")
    return Matrix(h.value)


def max_degree_vertex(A: Matrix):
    v, d = ctypes.c_uint32(), _S()
    check(lib().alp_hip_max_degree_vertex(A.h, ctypes.byref(v), ctypes.byref(d)), "max_degree_vertex")
    return v.value, d.value


def bfs(A: Matrix, source: int, out: "Vector" = None):
    """(levels as an int32 Vector, number of levels, vertices reached); `out`, an
    int32 Vector of A.nrows entries, receives the levels instead of a new one."""
    if out is not None and (out.n != A.nrows or out.dtype != np.int32):
        raise ValueError("bfs: out must be an int32 Vector of A.nrows entries")
    lev = Vector(A.nrows, np.int32) if out is None else out
    nl, reached = _S(), _S()
    check(lib().alp_hip_bfs(A.h, source, lev.ptr, ctypes.byref(nl), ctypes.byref(reached)), "bfs")
    return lev, nl.value, reached.value


def pagerank(L: Matrix, alpha=0.85, conv=1e-7, max_iter=1000):
    """(rc, ranks Vector, iterations, quality) as grb::algorithms::simple_pagerank."""
    pr = Vector(L.nrows)
    it, q = _S(), _D()
    rc = lib().alp_hip_pagerank(L.h, pr.ptr, alpha, conv, max_iter, ctypes.byref(it), ctypes.byref(q))
    return rc, pr, it.value, q.value


def khop(A: Matrix, source: int, k: int, out: "Vector" = None) -> Vector:
    """The k-hop set of source as a uint8 Vector (1: within k hops); `out`, a
    uint8 Vector of A.nrows entries, receives it instead of a new one."""
    if out is not None and (out.n != A.nrows or out.dtype != np.uint8):
        raise ValueError("khop: out must be a uint8 Vector of A.nrows entries")
    out = Vector(A.nrows, np.uint8) if out is None else out
    check(lib().alp_hip_khop(A.h, source, k, out.ptr), "khop")
    return out


# ---------------------------------------------------------------------------
# batched multi-RHS CG
# ---------------------------------------------------------------------------

class Batch:
    """Block-diagonal batch of SPD systems with nrhs right-hand sides: anisotropic
    7-point Laplacians (Batch(nx, ny, nz, cx, cy, cz)) or caller-supplied CRS
    systems (Batch.from_crs)."""

    def __init__(self, nx=None, ny=None, nz=None, cx=None, cy=None, cz=None, nrhs=8, _handle=None, _nsys=0):
        h = _P()
        if _handle is None:
            self.nsys = len(nx)
            arrs = [np.ascontiguousarray(a, np.uint32) for a in (nx, ny, nz)] + \
                   [np.ascontiguousarray(a, np.float64) for a in (cx, cy, cz)]
            check(lib().alp_hip_batch_create_laplace7(ctypes.byref(h), self.nsys, *[a.ctypes.data for a in arrs],
                                                      nrhs), "batch_create_laplace7")
            self.h = h.value
        else:
            self.nsys, self.h = _nsys, _handle
        self.nrhs = nrhs
        nr, nnz = _S(), _S()
        self.row0 = np.zeros(self.nsys + 1, np.uint64)
        check(lib().alp_hip_batch_info(self.h, ctypes.byref(nr), ctypes.byref(nnz), self.row0.ctypes.data),
              "batch_info")
        self.nrows, self.nnz = nr.value, nnz.value

    @classmethod
    def from_crs(cls, systems, nrhs=8) -> "Batch":
        """systems: a list of (off, idx, val) CRS arrays, one SPD matrix each
        (columns local to the system); stacked into one batch."""
        row0 = np.zeros(len(systems) + 1, np.uint64)
        offs, idxs, vals = [np.zeros(1, np.uint64)], [], []
        base = 0
        for s, (off, idx, val) in enumerate(systems):
            off = np.asarray(off, np.uint64)
            row0[s + 1] = row0[s] + (off.size - 1)
            offs.append(off[1:] + np.uint64(base))
            idxs.append(np.asarray(idx, np.uint32))
            vals.append(np.asarray(val, np.float64))
            base += int(off[-1])
        off = np.ascontiguousarray(np.concatenate(offs))
        idx = np.ascontiguousarray(np.concatenate(idxs)) if idxs else np.zeros(0, np.uint32)
        val = np.ascontiguousarray(np.concatenate(vals)) if vals else np.zeros(0)
        h = _P()
        check(lib().alp_hip_batch_create_crs(ctypes.byref(h), len(systems), row0.ctypes.data, off.ctypes.data,
                                             idx.ctypes.data, val.ctypes.data, nrhs), "batch_create_crs")
        return cls(nrhs=nrhs, _handle=h.value, _nsys=len(systems))

    def spmm(self, X: Vector, Y: Vector) -> None:
        check(lib().alp_hip_batch_spmm(self.h, X.ptr, Y.ptr), "batch_spmm")

    def layout(self):
        """(slices, slices in diagonal form) of the SpMM's diagonal layout (0 when not built)."""
        a, b = _S(), _S()
        check(lib().alp_hip_batch_layout(self.h, ctypes.byref(a), ctypes.byref(b)), "batch_layout")
        return a.value, b.value

    def autotune(self) -> int:
        w = _I()
        check(lib().alp_hip_batch_autotune(self.h, ctypes.byref(w)), "batch_autotune")
        return w.value

    def solve(self, B: Vector, X: Vector, tol: float, max_iter: int = 10000, precision: int = 0):
        """(rc, iterations per instance, true relative residual per instance, refinement steps)."""
        n = self.nsys * self.nrhs
        its = np.zeros(n, np.uint64)
        rel = np.zeros(n)
        outer = _S()
        rc = lib().alp_hip_batch_solve(self.h, B.ptr, X.ptr, tol, max_iter, precision, its.ctypes.data,
                                       rel.ctypes.data, ctypes.byref(outer))
        return rc, its.reshape(self.nsys, self.nrhs), rel.reshape(self.nsys, self.nrhs), outer.value

    def free(self) -> None:
        if getattr(self, "h", None):
            lib().alp_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
