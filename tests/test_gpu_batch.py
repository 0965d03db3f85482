"""Batched multi-RHS CG on the device (no counterpart in the reference: pinned
by an FP64 CPU restatement of the anisotropic 7-point Laplacian, direct solves on
small systems, and true FP64 residuals): SpMM within 1e-12 relative; every
(system, rhs) instance at ||b - A x|| <= tol ||b|| in FP64 and mixed precision."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def A():
    import alp_amd
    alp_amd.init(0)
    return alp_amd


def laplace7(nx, ny, nz, cx, cy, cz):
    """CPU FP64 CSR (scipy) of the anisotropic 7-point Laplacian, Dirichlet boundaries."""
    import scipy.sparse as sp
    def d1(n, c):
        return sp.diags([-c * np.ones(n - 1), 2 * c * np.ones(n), -c * np.ones(n - 1)], [-1, 0, 1])
    Ix, Iy, Iz = sp.identity(nx), sp.identity(ny), sp.identity(nz)
    A = (sp.kron(Iz, sp.kron(Iy, d1(nx, cx))) + sp.kron(Iz, sp.kron(d1(ny, cy), Ix)) +
         sp.kron(d1(nz, cz), sp.kron(Iy, Ix)))
    return A.tocsr()


def make_batch(A, rng, nsys, lo, hi, nrhs=8):
    nx, ny, nz = (rng.integers(lo, hi + 1, nsys) for _ in range(3))
    cx, cy, cz = (rng.lognormal(0.0, 0.5, nsys) for _ in range(3))
    return A.Batch(nx, ny, nz, cx, cy, cz, nrhs), (nx, ny, nz, cx, cy, cz)


def test_spmm_matches_cpu(A):
    rng = np.random.default_rng(4)
    B, p = make_batch(A, rng, 5, 3, 9)
    X = rng.standard_normal((B.nrows, 8))
    Y = A.Vector(B.nrows * 8)
    B.spmm(A.Vector.from_numpy(X.ravel()), Y)
    got = Y.numpy().reshape(B.nrows, 8)
    for s in range(B.nsys):
        r0, r1 = int(B.row0[s]), int(B.row0[s + 1])
        M = laplace7(*(int(v[s]) for v in p[:3]), *(float(v[s]) for v in p[3:]))
        want = M @ X[r0:r1]
        assert np.max(np.abs(got[r0:r1] - want)) <= 1e-12 * np.max(np.abs(want))
    assert B.nnz == sum(laplace7(*(int(v[s]) for v in p[:3]), 1, 1, 1).nnz for s in range(B.nsys))


def test_diagonal_form_equals_crs_bit_for_bit(A, monkeypatch):
    """The SpMM over the diagonal form (slices of <= 64 rows inside each tile; the
    slices that straddle a z-plane keep the CRS loop) folds each row in storage
    order and keeps the per-tile partials: the same batch built without the form
    (ALP_HIP_BATCH_DIA=0) gives bit-identical products and solves."""
    rng = np.random.default_rng(31)
    nx, ny, nz = (rng.integers(16, 41, 6) for _ in range(3))
    cx, cy, cz = (rng.lognormal(0.0, 0.5, 6) for _ in range(3))
    D = A.Batch(nx, ny, nz, cx, cy, cz, 8)
    nsl, ndia = D.layout()
    assert nsl > 0 and 0.5 * nsl <= ndia < nsl, (nsl, ndia)
    monkeypatch.setenv("ALP_HIP_BATCH_DIA", "0")
    C = A.Batch(nx, ny, nz, cx, cy, cz, 8)
    assert C.layout()[1] == 0
    monkeypatch.delenv("ALP_HIP_BATCH_DIA")
    X = A.Vector.from_numpy(rng.standard_normal(D.nrows * 8))
    Yd, Yc = A.Vector(D.nrows * 8), A.Vector(D.nrows * 8)
    D.spmm(X, Yd)
    C.spmm(X, Yc)
    assert np.array_equal(Yd.numpy(), Yc.numpy())
    for s in range(D.nsys):
        r0, r1 = int(D.row0[s]), int(D.row0[s + 1])
        M = laplace7(int(nx[s]), int(ny[s]), int(nz[s]), float(cx[s]), float(cy[s]), float(cz[s]))
        want = M @ X.numpy().reshape(-1, 8)[r0:r1]
        assert np.max(np.abs(Yd.numpy().reshape(-1, 8)[r0:r1] - want)) <= 1e-12 * np.max(np.abs(want))
    # both at the default RHS width (no autotune): the same partial-sum trees
    rhs = A.Vector.from_numpy(rng.standard_normal(D.nrows * 8))
    for precision in (0, 1):
        Xd, Xc = A.Vector(D.nrows * 8, fill=0.0), A.Vector(D.nrows * 8, fill=0.0)
        rd = D.solve(rhs, Xd, 1e-8, 5000, precision)
        rc = C.solve(rhs, Xc, 1e-8, 5000, precision)
        assert rd[0] == A.SUCCESS and rc[0] == A.SUCCESS
        assert np.array_equal(rd[1], rc[1]) and np.array_equal(Xd.numpy(), Xc.numpy()), precision
        assert np.all(rd[2] <= 1e-8)


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_small_batch_against_direct_solves(A, precision):
    import scipy.sparse.linalg as spla
    rng = np.random.default_rng(7 + precision)
    B, p = make_batch(A, rng, 7, 4, 12)
    assert B.autotune() in (1, 2, 4, 8)
    rhs = rng.standard_normal((B.nrows, 8))
    X = A.Vector(B.nrows * 8, fill=0.0)
    rc, its, rel, outer = B.solve(A.Vector.from_numpy(rhs.ravel()), X, 1e-8, 5000, precision)
    assert rc == A.SUCCESS
    assert np.all(rel <= 1e-8)
    x = X.numpy().reshape(B.nrows, 8)
    for s in range(B.nsys):
        r0, r1 = int(B.row0[s]), int(B.row0[s + 1])
        M = laplace7(*(int(v[s]) for v in p[:3]), *(float(v[s]) for v in p[3:]))
        res = np.linalg.norm(rhs[r0:r1] - M @ x[r0:r1], axis=0) / np.linalg.norm(rhs[r0:r1], axis=0)
        assert np.all(res <= 1.01e-8)
        xs = spla.spsolve(M.tocsc(), rhs[r0:r1])
        assert np.max(np.abs(x[r0:r1] - xs)) <= 1e-6 * np.max(np.abs(xs))


def test_c4_batch_64_systems(A):
    rng = np.random.default_rng(2024)
    B, p = make_batch(A, rng, 64, 16, 96)
    B.autotune()
    rhs = rng.standard_normal(B.nrows * 8)
    maxits = {}
    for precision in (0, 1, 2):
        X = A.Vector(B.nrows * 8, fill=0.0)
        rc, its, rel, outer = B.solve(A.Vector.from_numpy(rhs), X, 1e-8, 20000, precision)
        assert rc == A.SUCCESS, precision
        assert np.all(rel <= 1e-8) and its.min() > 0
        maxits[precision] = int(its.max())
        # spot-check the device's own residual report on three systems on the CPU
        x = X.numpy().reshape(B.nrows, 8)
        r = rhs.reshape(B.nrows, 8)
        for s in (0, 31, 63):
            r0, r1 = int(B.row0[s]), int(B.row0[s + 1])
            M = laplace7(*(int(v[s]) for v in p[:3]), *(float(v[s]) for v in p[3:]))
            res = np.linalg.norm(r[r0:r1] - M @ x[r0:r1], axis=0) / np.linalg.norm(r[r0:r1], axis=0)
            assert np.all(res <= 1.01e-8)
            assert np.allclose(res, rel[s], rtol=1e-3, atol=1e-12)
    # reliable updates keep each instance's Krylov space: near the FP64 count
    assert maxits[1] <= 1.3 * maxits[0], maxits


def test_batch_reliable_updates_stop_at_the_iteration_limit(A):
    rng = np.random.default_rng(11)
    B, p = make_batch(A, rng, 5, 12, 20)
    rhs = rng.standard_normal(B.nrows * 8)
    X = A.Vector(B.nrows * 8, fill=0.0)
    rc, its, rel, outer = B.solve(A.Vector.from_numpy(rhs), X, 1e-12, 6, 1)
    assert rc == A.FAILED
    assert its.max() == 6
    assert np.all(np.isfinite(rel)) and np.all(rel > 1e-12)


def random_spd(n, deg, rng):
    """Uneven SPD system: a random symmetric pattern with weights in (-1, 0] and a
    diagonal of the row's absolute sum plus a random shift in [0.01, 1) (strictly
    diagonally dominant), columns ascending in each row."""
    import scipy.sparse as sp
    i = rng.integers(0, n, n * deg)
    j = rng.integers(0, n, n * deg)
    keep = i != j
    w = -rng.random(keep.sum())
    M = sp.coo_matrix((w, (i[keep], j[keep])), shape=(n, n)).tocsr()
    M = M + M.T
    M = M + sp.diags(-np.asarray(M.sum(axis=1)).ravel() + rng.uniform(0.01, 1.0, n))
    M = M.tocsr()
    M.sum_duplicates()
    M.sort_indices()
    return M


def test_crs_batch_equals_generated_laplacians_bit_for_bit(A):
    """The same anisotropic Laplacians given as caller CRS (columns ascending, as
    the generator stores them) give the generator's batch: identical products and
    solves, bit for bit, in both precisions."""
    rng = np.random.default_rng(5)
    G, p = make_batch(A, rng, 6, 8, 24)
    mats = []
    for s in range(G.nsys):
        M = laplace7(*(int(v[s]) for v in p[:3]), *(float(v[s]) for v in p[3:]))
        M.sort_indices()
        mats.append((M.indptr, M.indices, M.data))
    C = A.Batch.from_crs(mats, 8)
    assert C.nrows == G.nrows and C.nnz == G.nnz and np.array_equal(C.row0, G.row0)
    assert C.layout() == G.layout()
    X = A.Vector.from_numpy(rng.standard_normal(G.nrows * 8))
    Yg, Yc = A.Vector(G.nrows * 8), A.Vector(G.nrows * 8)
    G.spmm(X, Yg)
    C.spmm(X, Yc)
    assert np.array_equal(Yg.numpy(), Yc.numpy())
    rhs = A.Vector.from_numpy(rng.standard_normal(G.nrows * 8))
    for precision in (0, 1):
        Xg, Xc = A.Vector(G.nrows * 8, fill=0.0), A.Vector(G.nrows * 8, fill=0.0)
        rg = G.solve(rhs, Xg, 1e-8, 5000, precision)
        rc = C.solve(rhs, Xc, 1e-8, 5000, precision)
        assert rg[0] == rc[0] == A.SUCCESS
        assert np.array_equal(rg[1], rc[1]) and np.array_equal(Xg.numpy(), Xc.numpy()), precision


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_uneven_crs_systems_against_direct_solves(A, precision):
    """Caller-supplied systems of uneven size and row length (random SPD patterns,
    40 to 3000 rows, 1 to 40 entries per row): every instance reaches the FP64
    relative residual 1e-8 and agrees with scipy's direct solve."""
    import scipy.sparse.linalg as spla
    rng = np.random.default_rng(100 + precision)
    sizes = [40, 3000, 517, 1200, 64, 2049, 255, 900]
    mats = [random_spd(n, int(rng.integers(1, 20)), rng) for n in sizes]
    B = A.Batch.from_crs([(M.indptr, M.indices, M.data) for M in mats], 8)
    assert B.nsys == len(sizes) and B.nrows == sum(sizes)
    assert B.autotune() in (1, 2, 4, 8)
    rhs = rng.standard_normal((B.nrows, 8))
    X = A.Vector(B.nrows * 8, fill=0.0)
    rc, its, rel, outer = B.solve(A.Vector.from_numpy(rhs.ravel()), X, 1e-8, 20000, precision)
    assert rc == A.SUCCESS
    assert np.all(rel <= 1e-8)
    x = X.numpy().reshape(B.nrows, 8)
    for s, M in enumerate(mats):
        r0, r1 = int(B.row0[s]), int(B.row0[s + 1])
        res = np.linalg.norm(rhs[r0:r1] - M @ x[r0:r1], axis=0) / np.linalg.norm(rhs[r0:r1], axis=0)
        assert np.all(res <= 1.01e-8)
        xs = spla.spsolve(M.tocsc(), rhs[r0:r1])
        assert np.max(np.abs(x[r0:r1] - xs)) <= 1e-6 * np.max(np.abs(xs))


def test_crs_batch_rejects_malformed_input(A):
    import ctypes
    lib = A.lib()
    h = ctypes.c_void_p()
    row0 = np.array([0, 2], np.uint64)
    off = np.array([0, 1, 2], np.uint64)
    val = np.array([1.0, 1.0])
    bad_col = np.array([0, 2], np.uint32)  # column 2 outside a 2-row system
    assert lib.alp_hip_batch_create_crs(ctypes.byref(h), 1, row0.ctypes.data, off.ctypes.data,
                                        bad_col.ctypes.data, val.ctypes.data, 8) == A.MISMATCH
    ok = np.array([0, 1], np.uint32)
    bad_off = np.array([0, 2, 1], np.uint64)
    assert lib.alp_hip_batch_create_crs(ctypes.byref(h), 1, row0.ctypes.data, bad_off.ctypes.data,
                                        ok.ctypes.data, val.ctypes.data, 8) == A.ILLEGAL
    assert lib.alp_hip_batch_create_crs(ctypes.byref(h), 1, row0.ctypes.data, off.ctypes.data,
                                        ok.ctypes.data, val.ctypes.data, 3) == A.ILLEGAL  # nrhs
    empty = np.array([0, 0], np.uint64)
    assert lib.alp_hip_batch_create_crs(ctypes.byref(h), 1, empty.ctypes.data, off.ctypes.data,
                                        ok.ctypes.data, val.ctypes.data, 8) == A.ILLEGAL


def test_batch_rejects_unaligned_block_vectors(A):
    """The SpMM moves each row's right-hand sides in 16-byte chunks: block vectors
    must be 16-byte aligned (every allocation is); an offset pointer is refused."""
    rng = np.random.default_rng(3)
    B, _ = make_batch(A, rng, 2, 4, 6)
    X = A.Vector(B.nrows * 8 + 2, fill=1.0)
    Y = A.Vector(B.nrows * 8 + 2)
    lib = A.lib()
    assert lib.alp_hip_batch_spmm(B.h, X.ptr + 8, Y.ptr) == A.ILLEGAL
    assert lib.alp_hip_batch_spmm(B.h, X.ptr, Y.ptr + 8) == A.ILLEGAL
    assert lib.alp_hip_batch_spmm(B.h, X.ptr + 16, Y.ptr + 16) == A.SUCCESS
