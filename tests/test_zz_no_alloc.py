"""alp_hip_cg_noalloc through the C ABI against alp_hip_cg (precision 0): the
same kernels and reduction order, so x must agree bit for bit and the
iteration counts exactly. Two calls in a row, the second on a smaller system,
so the shared scratch array is reused with stale contents from the first.
A system above the scratch limit (2^20 rows) is refused with ALP_OUTOFMEM."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def alp():
    import alp_amd
    alp_amd.init(0)
    return alp_amd


def _random_spd_crs(n, deg, seed):
    """Symmetric, strictly diagonally dominant pattern matrix with random
    off-diagonal weights in (-1, 0): SPD, rows of uneven length."""
    rng = np.random.default_rng(seed)
    i = rng.integers(0, n, n * deg)
    j = rng.integers(0, n, n * deg)
    keep = i != j
    i, j = i[keep], j[keep]
    w = -rng.random(i.size) - 1e-3
    I = np.concatenate([i, j])
    J = np.concatenate([j, i])
    W = np.concatenate([w, w])
    key = I.astype(np.int64) * n + J
    key, first = np.unique(key, return_index=True)
    I, J, W = (key // n).astype(np.int64), (key % n).astype(np.int64), W[first]
    diag = np.zeros(n)
    np.add.at(diag, I, -W)
    I = np.concatenate([I, np.arange(n)])
    J = np.concatenate([J, np.arange(n)])
    W = np.concatenate([W, diag + 1.0])
    order = np.lexsort((J, I))
    I, J, W = I[order], J[order], W[order]
    off = np.zeros(n + 1, np.uint64)
    np.add.at(off, I + 1, 1)
    return np.cumsum(off).astype(np.uint64), J.astype(np.uint32), W


def _both(alp, A, b_host, x0_host, max_iter, tol):
    b = alp.Vector.from_numpy(b_host)
    x_ref = alp.Vector.from_numpy(x0_host)
    rc, it, res = alp.cg(A, b, x_ref, max_iter, tol, precision=0)
    x_na = alp.Vector.from_numpy(x0_host)
    it_na, res_na = ctypes.c_size_t(), ctypes.c_double()
    rc_na = alp.lib().alp_hip_cg_noalloc(A.h, b.ptr, x_na.ptr, max_iter, tol, ctypes.byref(it_na),
                                          ctypes.byref(res_na))
    return (rc, it, res, x_ref.numpy()), (rc_na, it_na.value, res_na.value, x_na.numpy())


def _assert_same(ref, na):
    rc, it, res, x = ref
    rc_na, it_na, res_na, x_na = na
    assert rc_na == rc
    assert it_na == it
    assert res_na == res
    assert np.array_equal(x_na, x)


def test_noalloc_cg_equals_cg_on_two_calls(alp):
    rng = np.random.default_rng(7)
    # first call: 48^3 stencil (110,592 rows), b = A 1, random x0
    A = alp.Matrix.stencil27(48, 48, 48)
    ones = alp.Vector(A.nrows, fill=1.0)
    b = alp.Vector(A.nrows, fill=0.0)
    alp.check(alp.mxv(b, A, ones), "mxv")
    print(f"noalloc: stencil 48^3 n={A.nrows}", flush=True)
    ref, na = _both(alp, A, b.numpy(), rng.random(A.nrows), 1000, 1e-10)
    assert ref[0] == alp.SUCCESS
    _assert_same(ref, na)
    # second call, smaller: an uneven random SPD matrix from CRS (20,000 rows)
    n = 20000
    off, idx, val = _random_spd_crs(n, 6, 11)
    B = alp.Matrix.from_crs(n, n, off, idx, val)
    print(f"noalloc: random SPD n={n} nnz={int(off[-1])}", flush=True)
    ref2, na2 = _both(alp, B, rng.standard_normal(n), np.zeros(n), 2000, 1e-10)
    assert ref2[0] == alp.SUCCESS
    _assert_same(ref2, na2)
    # an iteration limit that stops first: FAILED from both, same iterate
    ref3, na3 = _both(alp, B, rng.standard_normal(n), np.zeros(n), 5, 1e-10)
    assert ref3[0] == alp.FAILED
    _assert_same(ref3, na3)


def test_noalloc_cg_refuses_systems_above_the_scratch_limit(alp):
    A = alp.Matrix.stencil27(102, 102, 102)  # 1,061,208 rows > 2^20
    b = alp.Vector(A.nrows, fill=1.0)
    x = alp.Vector(A.nrows, fill=0.0)
    it, res = ctypes.c_size_t(), ctypes.c_double()
    rc = alp.lib().alp_hip_cg_noalloc(A.h, b.ptr, x.ptr, 10, 1e-10, ctypes.byref(it), ctypes.byref(res))
    assert rc == alp.OUTOFMEM
    assert np.all(x.numpy() == 0.0)
