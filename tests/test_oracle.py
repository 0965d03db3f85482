"""Pins the CPU oracle (oracle/alp_oracle.c) before anything is checked against it.

Each oracle function is checked against (1) the golden fixtures written from
ALP's own sequential reference backend (tests/golden/make_golden.py), bit-exact,
and (2) something the mathematics fixes independently of the reference:
closed forms, brute force on tiny inputs, or invariants, chosen so that a
dropped term, a wrong sign or a transposed operand fails.
"""
import numpy as np
import pytest

import oracle as O


# ---------------------------------------------------------------- storage ----

@pytest.mark.parametrize("sizes", [(3, 3, 3), (4, 5, 6)])
def test_stencil_storage_matches_reference(golden, sizes):
    g = golden("stencil_storage_%dx%dx%d" % sizes)
    A = O.stencil(*sizes)
    assert np.array_equal(A.off, g["crs_off"])
    assert np.array_equal(A.idx, g["crs_idx"])
    assert np.array_equal(A.val, g["crs_val"])
    # the stencil is symmetric, so the reference's CCS equals its CRS
    assert np.array_equal(g["ccs_idx"], g["crs_idx"])


def closed_form_A1(nx, ny, nz):
    """b_i = 27 - prod_d c_d(i), c_d = 2 on a face of dimension d, 3 inside (diag 26, off -1)."""
    def c(n):
        v = np.full(n, 3.0)
        v[0] = v[-1] = 2.0
        return v
    cx, cy, cz = c(nx), c(ny), c(nz)
    prod = cz[:, None, None] * cy[None, :, None] * cx[None, None, :]
    return (27.0 - prod).reshape(-1)


@pytest.mark.parametrize("sizes", [(3, 3, 3), (8, 8, 8), (5, 7, 9), (64, 64, 64)])
def test_stencil_closed_form(sizes):
    A = O.stencil(*sizes)
    b = O.spmv(A, np.ones(A.nrows))
    assert np.array_equal(b, closed_form_A1(*sizes))
    assert A.nnz == (3 * sizes[0] - 2) * (3 * sizes[1] - 2) * (3 * sizes[2] - 2)


def test_stencil_64_norm():
    # ||A*1||^2 at 64^3 as measured on the reference (interior 0, face 9, edge 15, corner 19)
    b = O.spmv(O.stencil(64, 64, 64), np.ones(64 ** 3))
    assert float(b @ b) == 2038472.0


def test_stencil_bruteforce_dense():
    nx, ny, nz = 3, 4, 5
    A = O.stencil(nx, ny, nz).dense()
    n = nx * ny * nz
    coords = [(i % nx, (i // nx) % ny, i // (nx * ny)) for i in range(n)]
    for i, ci in enumerate(coords):
        for j, cj in enumerate(coords):
            near = all(abs(a - b) <= 1 for a, b in zip(ci, cj))
            want = 26.0 if i == j else (-1.0 if near else 0.0)
            assert A[i, j] == want


def test_coo_to_compressed_reverse_order_and_dense():
    rng = np.random.default_rng(0)
    n, m = 13, 9
    key = rng.choice(n * m, 40, replace=False)
    I, J = (key // m).astype(np.uint32), (key % m).astype(np.uint32)
    V = rng.standard_normal(40)
    crs = O.coo_to_compressed(n, m, I, J, V, which=0)
    ccs = O.coo_to_compressed(n, m, I, J, V, which=1)
    D = np.zeros((n, m))
    D[I, J] = V
    assert np.array_equal(crs.dense(), D)
    assert np.array_equal(ccs.dense(), D.T)
    # within a row, entries appear in reverse input order (matrix.hpp:1918-1930)
    for i in range(n):
        inp = [J[k] for k in range(40) if I[k] == i]
        assert list(crs.idx[crs.off[i]:crs.off[i + 1]]) == inp[::-1]


# ------------------------------------------------------------------- SpMV ----

@pytest.mark.parametrize("sizes", [(8, 8, 8), (5, 9, 7), (16, 16, 16)])
def test_spmv_matches_reference(golden, sizes):
    g = golden("mxv_%dx%dx%d" % sizes)
    A = O.stencil(*sizes)
    assert np.array_equal(O.spmv(A, g["x"], g["y0"]), g["y"])
    # A is symmetric: the reference's transposed (CCS) product gives the same bits
    assert np.array_equal(O.spmv(A, g["x"], g["y0"]), g["yt"])


def test_spmv_dense_bruteforce_tolerance():
    rng = np.random.default_rng(1)
    A = O.stencil(6, 5, 4)
    x = rng.standard_normal(A.nrows)
    y0 = rng.standard_normal(A.nrows)
    ref = y0 + A.dense() @ x
    assert np.allclose(O.spmv(A, x, y0), ref, rtol=0, atol=1e-12)
    # an off-by-one column index would not pass: perturb and check detection
    bad = O.CSR(A.nrows, A.ncols, A.off, np.minimum(A.idx + 1, A.ncols - 1), A.val)
    assert not np.allclose(O.spmv(bad, x, y0), ref, rtol=0, atol=1e-6)


def test_spmv_f32_matches_rounded_f64_within_ulps():
    rng = np.random.default_rng(2)
    A = O.stencil(8, 8, 8)
    x = rng.random(A.nrows).astype(np.float32)
    y32 = O.spmv(A, x)
    y64 = O.spmv(A, x.astype(np.float64))
    assert np.allclose(y32, y64, rtol=1e-5, atol=1e-5)


def test_lorland_spmv_bruteforce():
    rng = np.random.default_rng(3)
    n = 50
    D = (rng.random((n, n)) < 0.08)
    I, J = np.nonzero(D)
    A = O.coo_to_compressed(n, n, I, J, which=0)
    x = (rng.random(n) < 0.3).astype(np.uint8)
    y0 = (rng.random(n) < 0.2).astype(np.uint8)
    want = y0 | (D.astype(np.uint8) @ x > 0)
    assert np.array_equal(O.spmv_lorland(A, x, y0), want.astype(np.uint8))
    mask = (rng.random(n) < 0.5).astype(np.uint8)
    got = O.spmv_lorland(A, x, y0, mask=mask, invert_mask=True)
    want2 = np.where(mask == 0, want, y0)
    assert np.array_equal(got, want2.astype(np.uint8))


# -------------------------------------------------------------------- dot ----

def test_dot_matches_reference(golden):
    g = golden("dot")
    for n in (1, 2, 9, 10, 17, 1000, 4097):
        assert O.dot(g[f"x_{n}"], g[f"y_{n}"]) == g[f"z_{n}"][0], n


def test_dot_exact_integers_and_closed_forms():
    for n in (0, 1, 7, 8, 9, 16, 17, 1000):
        x = np.arange(n, dtype=np.float64)
        assert O.dot(x, np.ones(n)) == n * (n - 1) / 2
        assert O.dot(x, x) == sum(i * i for i in range(n))
    # the reference's dot.cpp closed form: x = 1.5, y = -1 except y[0] = 2.55 -> 2.55 - 1.5(n-1)... sign pinned
    n = 1024
    x = np.full(n, 1.5)
    y = -np.ones(n)
    assert O.dot(x, y) == -1.5 * n


# --------------------------------------------------------------------- CG ----

@pytest.mark.parametrize("name,sizes", [("cg_8x8x8_it1000", (8, 8, 8)),
                                        ("cg_12x10x9_it1000", (12, 10, 9)),
                                        ("cg_16x16x16_it20", (16, 16, 16))])
def test_cg_matches_reference(golden, name, sizes):
    g = golden(name)
    A = O.stencil(*sizes)
    b = O.spmv(A, np.ones(A.nrows))
    max_iter = int(name.split("_it")[1])
    rc, x, it, res, _ = O.cg(A, b, g["x0"], max_iter, 1e-10)
    assert it == g["iterations"][0]
    assert res == g["residual"][0]
    assert np.array_equal(x, g["x"])
    assert (rc == 0) == (g["rc"][0] == 0)


def test_cg_64_50_iterations_matches_reference(golden):
    import hashlib
    g = golden("cg_64x64x64_it50")
    A = O.stencil(64, 64, 64)
    b = O.spmv(A, np.ones(A.nrows))
    rc, x, it, res, _ = O.cg(A, b, g["x0"] if "x0" in g else np.random.default_rng(42).random(A.nrows),
                             50, 1e-300)
    assert it == 50 and rc == 1
    assert res == g["residual"][0]
    assert np.array_equal(x[g["x_probe_idx"]], g["x_probe"])
    assert hashlib.sha256(x.tobytes()).digest() == bytes(g["x_sha256"])


def test_cg_64_to_1e10_matches_reference(golden):
    """The oracle's CG to a 1e-10 relative residual at 64^3 (x0 = 0, b = A*1):
    the reference's iteration count, return code and residual exactly, and x at
    the probed entries bit for bit (the fixture's 192^3 count is checked on the
    device only)."""
    g = golden("cg_64x64x64_tol1e-10")
    A = O.stencil(64, 64, 64)
    b = O.spmv(A, np.ones(A.nrows))
    rc, x, it, res, _ = O.cg(A, b, np.zeros(A.nrows), 1000, 1e-10)
    assert rc == 0 == int(g["rc"][0])
    assert it == int(g["iterations"][0])
    assert res == g["residual"][0]
    assert np.array_equal(x[g["x_probe_idx"]], g["x_probe"])


def test_cg_against_direct_solve():
    A = O.stencil(5, 6, 7)
    rng = np.random.default_rng(5)
    b = rng.standard_normal(A.nrows)
    rc, x, it, res, hist = O.cg(A, b, np.zeros(A.nrows), 500, 1e-12)
    assert rc == 0
    xs = np.linalg.solve(A.dense(), b)
    assert np.max(np.abs(x - xs)) < 1e-10 * np.max(np.abs(xs))
    # the returned residual is ||r||^2 of the recursively updated r
    assert res == hist[-1] and res < (1e-12 ** 2) * float(b @ b)


# ------------------------------------------------------------------- HPCG ----

@pytest.mark.parametrize("name,sizes,levels,colors", [
    ("hpcg_16x16x16_l1_c27", (16, 16, 16), 1, 27),
    ("hpcg_16x16x16_l3_c8", (16, 16, 16), 3, 8),
    ("hpcg_8x12x16_l2_c5", (8, 12, 16), 2, 5),
    ("hpcg_32x32x32_l3_c8", (32, 32, 32), 3, 8),
])
def test_hpcg_matches_reference(golden, name, sizes, levels, colors):
    import json
    import os
    meta = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "manifest.json")))[name]
    g = golden(name)
    cfg = O.hpcg_cfg(*sizes, levels, colors=colors)
    assert cfg.levels == meta["coarse_levels_built"] + 1
    A = O.stencil(*sizes)
    b = O.spmv(A, np.ones(A.nrows))
    x, it, res, _ = O.hpcg(cfg, b, np.zeros(A.nrows), meta["max_iter"], meta["tol"])
    assert it == g["iterations"][0]
    assert res == g["residual"][0]
    assert np.array_equal(x, g["x"])


def test_hpcg_smoke_criterion():
    # tests/smoke/hpcg.cpp:93: ||x - 1|| < 4e-14 after 56 iterations at 16^3, 27 colours
    A = O.stencil(16, 16, 16)
    b = O.spmv(A, np.ones(A.nrows))
    x, it, res, _ = O.hpcg(O.hpcg_cfg(16, 16, 16, 1, colors=27), b, np.zeros(A.nrows), 56, 0.0)
    assert np.linalg.norm(x - 1) < 4e-14


def test_rbgs_geometric_colouring_is_proper_and_converges():
    cfg = O.hpcg_cfg(16, 16, 16, 3, geometric=True)
    A = O.stencil(16, 16, 16)
    b = O.spmv(A, np.ones(A.nrows))
    x, it, res, hist = O.hpcg(cfg, b, np.zeros(A.nrows), 100, 1e-10)
    assert res / np.linalg.norm(b) <= 1e-10 * 1.01
    assert np.max(np.abs(x - 1)) < 1e-8
    # parity colouring: no stencil neighbour shares a colour
    n = A.nrows
    col = np.array([(i % 16 & 1) + 2 * ((i // 16) % 16 & 1) + 4 * ((i // 256) & 1) for i in range(n)])
    rows = np.repeat(np.arange(n), np.diff(A.off.astype(np.int64)))
    off_diag = rows != A.idx
    assert not np.any(col[rows[off_diag]] == col[A.idx[off_diag]])


def test_mg_vcycle_is_symmetric_preconditioner():
    # a symmetric V-cycle (forward then backward sweeps, R^T prolongation) is a
    # symmetric operator: <M r1, r2> == <r1, M r2>
    cfg = O.hpcg_cfg(8, 8, 8, 2, geometric=True)
    rng = np.random.default_rng(9)
    r1, r2 = rng.standard_normal(512), rng.standard_normal(512)
    a = O.mg_vcycle(cfg, r1) @ r2
    b = r1 @ O.mg_vcycle(cfg, r2)
    assert abs(a - b) < 1e-12 * abs(a)


# ----------------------------------------------------------------- graphs ----

def test_pagerank_smoke_vector(golden):
    g = golden("pagerank_smoke10")
    n = 10
    crs = O.coo_to_compressed(n, n, g["I"], g["J"], which=0)
    ccs = O.coo_to_compressed(n, n, g["I"], g["J"], which=1)
    rc, pr, it, q = O.pagerank(crs, ccs, 0.85, 1e-7, 1000)
    assert rc == 0 and it == g["iterations"][0]
    assert np.array_equal(pr, g["pr"])
    # the printed vector of tests/smoke/smoketests.sh:154 (6 significant digits)
    want = [0.106896, 0.105862, 0.104983, 0.104235, 0.1036, 0.10306, 0.102601, 0.102211,
            0.0584396, 0.108113]
    assert np.allclose(pr, want, rtol=5e-6, atol=0)


@pytest.mark.parametrize("scale", [9, 11])
def test_graph_fixtures(golden, scale):
    g = golden(f"graph_rmat_s{scale}")
    n = 1 << scale
    crs = O.coo_to_compressed(n, n, g["I"], g["J"], which=0)
    ccs = O.coo_to_compressed(n, n, g["I"], g["J"], which=1)
    rc, pr, it, q = O.pagerank(crs, ccs, 0.85, 1e-9, 1000)
    assert it == g["iterations"][0] and q == g["quality"][0]
    assert np.array_equal(pr, g["pr"])
    # PageRank invariant: entries positive and sum to 1 up to rounding
    assert abs(pr.sum() - 1.0) < 1e-9 and pr.min() > 0
    src = int(g["source"][0])
    lev = O.bfs_levels(crs, src)
    for k in (1, 2, 3):
        assert np.array_equal(O.khop(ccs, src, k), g[f"knn_{k}"])
        assert np.array_equal(((lev >= 0) & (lev <= k)).astype(np.uint8), g[f"knn_{k}"])


def test_bfs_levels_bruteforce():
    rng = np.random.default_rng(4)
    n = 80
    s, d = rng.integers(0, n, 160), rng.integers(0, n, 160)
    I, J = O.symmetrize_dedup(n, s.astype(np.uint32), d.astype(np.uint32))
    A = O.coo_to_compressed(n, n, I, J, which=0)
    D = A.dense() > 0
    lev = O.bfs_levels(A, 0)
    reach = np.zeros(n, bool)
    reach[0] = True
    want = np.full(n, -1)
    want[0] = 0
    for h in range(1, n):
        nxt = reach | (D.astype(int) @ reach.astype(int) > 0)
        want[nxt & ~reach] = h
        if (nxt == reach).all():
            break
        reach = nxt
    assert np.array_equal(lev, want)


def test_rmat_generator_properties():
    s, d = O.rmat_edges(10, 200000, seed=1)
    s2, d2 = O.rmat_edges(10, 200000, seed=1)
    assert np.array_equal(s, s2) and np.array_equal(d, d2)
    assert s.max() < 1024 and d.max() < 1024
    # top-level quadrant frequencies: a = .57, b = .19, c = .19, d = .05
    top_s, top_d = s >> 9, d >> 9
    f = [np.mean((top_s == i) & (top_d == j)) for i, j in ((0, 0), (0, 1), (1, 0), (1, 1))]
    assert np.allclose(f, [0.57, 0.19, 0.19, 0.05], atol=0.005)
    s3, _ = O.rmat_edges(10, 1000, seed=2)
    assert not np.array_equal(s3, s[:1000])


# ---- transition CG C ABI (src/transition/solver.cpp) --------------------------

TRANSITION_CASES = ["d_plain", "d_jacobi", "d_maxit", "d_precond_fails", "d_jacobi_large", "s_plain", "s_jacobi"]


def transition_case(golden, name):
    z = golden("transition_cg")
    g = {k[len(name) + 1:]: z[k] for k in z.files if k.startswith(name + "_")}
    meta = g["meta"]
    return g, dict(rc=int(meta[0]), iters=int(meta[1]), max_iter=int(meta[2]), fail_at=int(meta[3]),
                   precond=bool(meta[4]))


def jacobi_for(g, fail_at, dtype):
    ia, ja, a = g["ia"].astype(np.int64), g["ja"].astype(np.int64), g["a"]
    n = len(ia) - 1
    rows = np.repeat(np.arange(n), np.diff(ia))
    diag = np.zeros(n, dtype)
    diag[rows[ja == rows]] = a[ja == rows]
    calls = [0]

    def fn(z, r):
        calls[0] += 1
        if fail_at and calls[0] == fail_at:
            return 7
        np.divide(r, diag, out=z)
        return 0

    return fn, calls


@pytest.mark.parametrize("name", TRANSITION_CASES)
def test_oracle_pcg_matches_reference_transition_solver(golden, name):
    """The oracle's (preconditioned) CG reproduces ALP's transition solver bit for
    bit: same return code, iteration count, residual and x, in FP64 and FP32."""
    g, m = transition_case(golden, name)
    dt = g["a"].dtype
    n = len(g["ia"]) - 1
    cb = None
    if m["precond"]:
        fn, _ = jacobi_for(g, m["fail_at"], dt)
        cb = O.precond_callback(fn, n, dt)
    status, x, iters, res = O.pcg(g["ia"], g["ja"], g["a"], g["b"], g["x0"], m["max_iter"],
                                  float(g["tol"][0]), cb)
    assert (0 if status == 0 else 4) == m["rc"]  # NO_ERROR / FAILED
    assert iters == m["iters"]
    assert np.array_equal(x, g["x"])
    assert res == g["residual"][0]


def test_oracle_pcg_jacobi_on_constant_diagonal_is_plain_cg():
    """Closed-form check: Jacobi with a constant diagonal only rescales z, so the
    iterates equal the unpreconditioned ones up to rounding."""
    A = O.stencil(5, 4, 3)
    n = len(A.off) - 1
    b = np.random.default_rng(2).standard_normal(n)
    s0, x0, i0, r0 = O.pcg(A.off, A.idx, A.val, b, np.zeros(n), 1000, 1e-10)
    cb = O.precond_callback(lambda z, r: np.divide(r, 26.0, out=z), n)
    s1, x1, i1, r1 = O.pcg(A.off, A.idx, A.val, b, np.zeros(n), 1000, 1e-10, cb)
    assert s0 == s1 == 0 and i0 == i1
    assert np.max(np.abs(x0 - x1)) <= 1e-12 * np.max(np.abs(x0))


def test_reference_timing_entry_points():
    """The bench's CPU reference legs: the compiled reference's CG and HPCG run
    the requested number of iterations and report a positive time."""
    import ctypes
    if not O.ref_available("omp"):
        pytest.skip("oracle/_ref (the compiled reference) is not built here")
    R = O.ref("omp")
    it = ctypes.c_size_t()
    assert R.ref_time_stencil_cg(16, 16, 16, 5, 1e-300, ctypes.byref(it)) > 0 and it.value == 5
    assert R.ref_time_hpcg(16, 16, 16, 2, 8, 3, ctypes.byref(it)) > 0 and it.value == 3


# ------------------------------------------------------------------- mxm ----

def mxm_case(g, name):
    """The fixture's operands as the reference stored them (buildMatrixUnique
    order; a transposed operand by its CCS, the rows of its transpose)."""
    m, k, n, ring, tr, inplace, seed = (int(v) for v in g[f"{name}_meta"])
    vals = ring != 2
    ar, ac = (k, m) if tr == 1 else (m, k)
    br, bc = (n, k) if tr == 2 else (k, n)
    A = O.coo_to_compressed(ar, ac, g[f"{name}_IA"], g[f"{name}_JA"], g[f"{name}_VA"] if vals else None,
                            which=1 if tr == 1 else 0)
    B = O.coo_to_compressed(br, bc, g[f"{name}_IB"], g[f"{name}_JB"], g[f"{name}_VB"] if vals else None,
                            which=1 if tr == 2 else 0)
    C = O.coo_to_compressed(m, n, g[f"{name}_IC"], g[f"{name}_JC"], g[f"{name}_VC"]) if inplace else None
    return A, B, C, (0 if ring == 2 else ring)


def ccs_ascending(C):
    """The CCS of a CRS with each column's rows ascending (how mxm_generic fills
    an output's CCS, reference/blas3.hpp:1622-1636)."""
    rows = np.repeat(np.arange(C.nrows), np.diff(C.off.astype(np.int64)))
    order = np.argsort(C.idx, kind="stable")
    off = np.zeros(C.ncols + 1, np.uint64)
    off[1:] = np.cumsum(np.bincount(C.idx.astype(np.int64), minlength=C.ncols))
    return off, rows[order].astype(np.uint32), (None if C.val is None else C.val[order])


MXM_CASES = ["plus_times", "plus_times_inplace", "min_plus", "min_plus_inplace", "pattern",
             "transpose_left", "transpose_right", "wide"]


@pytest.mark.parametrize("name", MXM_CASES)
def test_mxm_matches_reference(golden, name):
    """orc_mxm against ALP's grb::mxm: the CRS (offsets, column order, values) bit
    for bit, and the reference's CCS is the ascending-row transpose of it."""
    g = golden("mxm")
    A, B, C, ring = mxm_case(g, name)
    got = O.mxm(A, B, ring, C)
    assert np.array_equal(got.off, g[f"{name}_crs_off"])
    assert np.array_equal(got.idx, g[f"{name}_crs_idx"])
    assert np.array_equal(got.val, g[f"{name}_crs_val"])
    coff, cidx, cval = ccs_ascending(got)
    assert np.array_equal(coff, g[f"{name}_ccs_off"]) and np.array_equal(cidx, g[f"{name}_ccs_idx"])
    assert np.array_equal(cval, g[f"{name}_ccs_val"])


@pytest.mark.parametrize("ring", [0, 1])
def test_mxm_dense_bruteforce(ring):
    """Values against a dense product independent of any order: plus-times within
    rounding (1e-13 relative), min-plus exactly; the pattern is that of the dense
    product's reachable entries."""
    rng = np.random.default_rng(11 + ring)
    m, k, n = 23, 19, 27
    def rand(r, c):
        D = np.where(rng.random((r, c)) < 0.2, rng.standard_normal((r, c)), 0.0)
        I, J = np.nonzero(D)
        return O.coo_to_compressed(r, c, I, J, D[I, J]), (D != 0)
    A, Ma = rand(m, k)
    B, Mb = rand(k, n)
    C = O.mxm(A, B, ring)
    Ad, Bd = A.dense(), B.dense()
    reach = (Ma.astype(int) @ Mb.astype(int)) > 0
    got = np.zeros((m, n))
    has = np.zeros((m, n), bool)
    for i in range(m):
        for q in range(int(C.off[i]), int(C.off[i + 1])):
            got[i, C.idx[q]] = C.val[q]
            has[i, C.idx[q]] = True
    assert np.array_equal(has, reach)
    if ring == 0:
        want = Ad @ Bd
        assert np.max(np.abs(got - want)) <= 1e-13 * max(1.0, np.max(np.abs(want)))
    else:
        S = np.where(Ma[:, :, None] & Mb[None, :, :], Ad[:, :, None] + Bd[None, :, :], np.inf).min(axis=1)
        assert np.array_equal(got[reach], S[reach])


def test_mxm_shifted_identities_closed_form():
    """The reference's unit test (tests/unit/mxm.cpp): eye(n, n, 1, 1) times
    eye(n, n, 2, 2) is eye(n, n, 2, 3); updating that result in place by the same
    product doubles it."""
    n = 10
    def eye(v, o):
        I = np.arange(n - o, dtype=np.uint32)
        return O.coo_to_compressed(n, n, I, I + o, np.full(n - o, float(v)))
    C = O.mxm(eye(1, 1), eye(2, 2))
    E = eye(2, 3)
    assert np.array_equal(C.off, E.off) and np.array_equal(C.idx, E.idx) and np.array_equal(C.val, E.val)
    C2 = O.mxm(eye(1, 1), eye(2, 2), 0, C)
    assert np.array_equal(C2.idx, E.idx) and np.array_equal(C2.val, 2 * E.val)


def test_mpv_fixture_closed_form(golden):
    """The mpv fixture (ALP's algorithms::mpv over plus-times, written from the
    reference by make_golden.py) against what its loop computes in closed form:
    each product is folded into the vector it writes and only temp is cleared
    once, so k = 0..5 give v, Av, A^2 v, Av + A^3 v, 2A^2 v + A^4 v and
    Av + 3A^3 v + A^5 v (dense products here; the fixture's unassigned entries
    are the rows the products never reach)."""
    g = golden("mpv")
    n = len(g["v"])
    A = np.zeros((n, n))
    A[g["I"], g["J"]] = g["V"]
    v = g["v"]
    p = [v]
    for _ in range(5):
        p.append(A @ p[-1])
    want = {0: p[0], 1: p[1], 2: p[2], 3: p[1] + p[3], 4: 2 * p[2] + p[4], 5: p[1] + 3 * p[3] + p[5]}
    reach = [np.ones(n, bool)]
    S = A != 0
    for _ in range(5):
        reach.append(S.astype(int) @ reach[-1].astype(int) > 0)
    assigned = {0: reach[0], 1: reach[1], 2: reach[2], 3: reach[1] | reach[3], 4: reach[2] | reach[4],
                5: reach[1] | reach[3] | reach[5]}
    for k in range(6):
        u = g[f"u_{k}"]
        assert np.array_equal(g[f"assigned_{k}"].astype(bool), assigned[k]), k
        assert np.max(np.abs(u - want[k])) <= 1e-13 * np.max(np.abs(want[k])), k
