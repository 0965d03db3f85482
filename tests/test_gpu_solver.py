"""The transition CG C ABI on the device (alp_hip_solver.h) against ALP's own
transition solver (src/transition/solver.cpp on the reference backend), via the
fixtures in tests/golden/transition_cg.npz: same return codes; the same
iteration counts within a small band (reductions run in a different order);
x within the tolerance the solve guarantees; exact counts where the solve
stops for a reason other than convergence."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = ["d_plain", "d_jacobi", "d_maxit", "d_precond_fails", "d_jacobi_large", "s_plain", "s_jacobi"]


@pytest.fixture(scope="module")
def S():
    import alp_amd
    alp_amd.init(0)
    from alp_amd import solver
    return solver


def load(golden, name):
    z = golden("transition_cg")
    g = {k[len(name) + 1:]: z[k] for k in z.files if k.startswith(name + "_")}
    m = g["meta"]
    return g, dict(rc=int(m[0]), iters=int(m[1]), max_iter=int(m[2]), fail_at=int(m[3]), precond=bool(m[4]))


def jacobi(g, fail_at=0):
    ia, ja, a = g["ia"].astype(np.int64), g["ja"].astype(np.int64), g["a"]
    n = len(ia) - 1
    rows = np.repeat(np.arange(n), np.diff(ia))
    diag = np.zeros(n, a.dtype)
    diag[rows[ja == rows]] = a[ja == rows]
    calls = [0]

    def fn(z, r):
        calls[0] += 1
        if fail_at and calls[0] == fail_at:
            return 7
        np.divide(r, diag, out=z)
        return 0

    return fn, calls


def true_residual(g, x):
    import scipy.sparse as sp
    n = len(g["ia"]) - 1
    A = sp.csr_matrix((g["a"].astype(np.float64), g["ja"].astype(np.int64), g["ia"].astype(np.int64)), shape=(n, n))
    b = g["b"].astype(np.float64)
    return np.linalg.norm(b - A @ x.astype(np.float64)) / np.linalg.norm(b)


def run(S, g, m, variant, fail_at=None):
    cg = S.SparseCG(variant, g["ia"], g["ja"], g["a"])
    cg.tolerance = float(g["tol"][0])
    cg.set_max_iter(m["max_iter"])
    calls = None
    if m["precond"]:
        fn, calls = jacobi(g, m["fail_at"] if fail_at is None else fail_at)
        assert cg.set_preconditioner(fn) == S.NO_ERROR
    x = g["x0"].copy()
    rc = cg.solve(x, g["b"])
    return cg, rc, x, calls


@pytest.mark.parametrize("name", CASES)
def test_matches_reference_transition_solver(S, golden, name):
    g, m = load(golden, name)
    f64 = g["a"].dtype == np.float64
    cg, rc, x, calls = run(S, g, m, "dii" if f64 else "sii")
    assert rc == m["rc"]
    tol = float(g["tol"][0])
    if m["rc"] == S.NO_ERROR:
        assert abs(cg.iterations - m["iters"]) <= (2 if f64 else 3)
        assert true_residual(g, x) <= (1.5 * tol if f64 else 4 * tol)
        ref = g["x"].astype(np.float64)
        assert np.max(np.abs(x - ref)) <= (1e-6 if f64 else 5e-3) * np.max(np.abs(ref))
        assert np.sqrt(cg.residual) < tol * np.linalg.norm(g["b"].astype(np.float64))
    else:
        # FAILED at the iteration limit or on the preconditioner's error: exact count
        assert cg.iterations == m["iters"]
        ref = g["x"].astype(np.float64)
        assert np.max(np.abs(x - ref)) <= 1e-9 * np.max(np.abs(ref))
        assert abs(cg.residual - g["residual"][0]) <= 1e-8 * g["residual"][0]
    if calls is not None:
        # one call in the prelude and one per iteration that continues
        expect = m["fail_at"] if m["fail_at"] else cg.iterations
        assert calls[0] == expect
    cg.close()


def test_index_variants_agree_bit_for_bit(S, golden):
    g, m = load(golden, "d_jacobi")
    xs = []
    for v in ("dii", "diz", "dzz"):
        cg, rc, x, _ = run(S, g, m, v)
        assert rc == S.NO_ERROR
        xs.append((x, cg.iterations, cg.residual))
        cg.close()
    for x, it, res in xs[1:]:
        assert np.array_equal(x, xs[0][0]) and it == xs[0][1] and res == xs[0][2]
    g, m = load(golden, "s_jacobi")
    a = run(S, g, m, "sii")[2]
    for v in ("siz", "szz"):
        assert np.array_equal(run(S, g, m, v)[2], a)


def test_handle_reuse_is_deterministic_and_keeps_settings(S, golden):
    g, m = load(golden, "d_plain")
    cg = S.SparseCG("dii", g["ia"], g["ja"], g["a"])
    assert cg.size == len(g["ia"]) - 1
    assert cg.tolerance == pytest.approx(1e-5) and cg.residual == np.inf and cg.iterations == 0
    cg.tolerance = 1e-10
    out = []
    for _ in range(2):
        x = g["x0"].copy()
        assert cg.solve(x, g["b"]) == S.NO_ERROR
        out.append((x, cg.iterations, cg.residual))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1:] == out[1][1:]
    # a converged initial guess still takes the one iteration of the do-while loop
    x = out[0][0].copy()
    assert cg.solve(x, g["b"]) == S.NO_ERROR and cg.iterations == 1
    cg.close()


def test_handle_of_another_variant_and_bad_settings(S, golden):
    g, m = load(golden, "d_plain")
    cg = S.SparseCG("dii", g["ia"], g["ja"], g["a"])
    x = g["x0"].copy()
    assert S.fn("solve", "dzz")(cg.h, x.ctypes.data, g["b"].ctypes.data) == S.ILLEGAL_ARGUMENT
    xs = g["x0"].astype(np.float32)
    assert S.fn("solve", "sii")(cg.h, xs.ctypes.data, xs.ctypes.data) == S.ILLEGAL_ARGUMENT
    # the reference CG rejects a non-positive tolerance / zero iterations: UNKNOWN
    cg.tolerance = 0.0
    assert cg.solve(x, g["b"]) == S.UNKNOWN
    cg.tolerance = 1e-10
    cg.set_max_iter(0)
    assert cg.solve(x, g["b"]) == S.UNKNOWN
    # NULL preconditioner with data is illegal; NULL/NULL removes it
    assert S.fn("set_preconditioner", "dii")(cg.h, None, 1234) == S.ILLEGAL_ARGUMENT
    assert cg.set_preconditioner(None) == S.NO_ERROR
    cg.close()


def test_stencil_64_cubed_against_native_cg(S):
    """A 262144-row system through the transition ABI reaches the same iteration
    count as the native device CG on the same matrix (within one)."""
    import alp_amd as A
    import oracle as O
    st = O.stencil(64, 64, 64)
    n = len(st.off) - 1
    ia = st.off.astype(np.uint64)
    ja = st.idx.astype(np.uint64)
    b = O.spmv(st, np.ones(n), np.zeros(n))
    cg = S.SparseCG("dzz", ia, ja, st.val)
    cg.tolerance = 1e-10
    x = np.zeros(n)
    assert cg.solve(x, b) == S.NO_ERROR
    assert np.max(np.abs(x - 1.0)) <= 1e-7
    M = A.Matrix.from_crs(n, n, st.off, st.idx, st.val)
    xv = A.Vector(n, fill=0.0)
    rc, it, res = A.cg(M, A.Vector.from_numpy(b), xv, 1000, 1e-10)
    assert rc == A.SUCCESS and abs(it - cg.iterations) <= 1
    cg.close()
