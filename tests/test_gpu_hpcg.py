"""Parity of the device HPCG (multigrid V-cycle, Gauss-Seidel colour sweeps,
injection transfers, preconditioned CG) against the oracle and the reference's
fixtures. Bars: V-cycle within 1e-12 relative of the oracle's; iteration counts
within +-1 of the reference's; solutions to the reference's / closed-form bar."""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def A():
    import alp_amd
    alp_amd.init(0)
    return alp_amd


def manifest(name):
    return json.load(open(os.path.join(HERE, "golden", "manifest.json")))[name]


def rhs_ones(A, nx, ny, nz):
    M = A.Matrix.stencil27(nx, ny, nz)
    b = A.Vector(M.nrows, fill=0.0)
    A.mxv(b, M, A.Vector(M.nrows, fill=1.0))
    return M, b


@pytest.mark.parametrize("sizes,levels,colours", [((8, 8, 8), 1, 0), ((16, 16, 16), 2, 0),
                                                   ((16, 16, 16), 2, 8), ((12, 8, 16), 1, 27),
                                                   ((8, 12, 16), 1, 5)])
def test_vcycle_matches_oracle(A, sizes, levels, colours):
    H = A.Hpcg(*sizes, max_levels=levels, colours=colours)
    geometric = colours == 0
    assert H.proper == geometric  # parity colouring is proper, index mod k is not
    cfg = O.hpcg_cfg(*sizes, levels, colors=colours if colours else 8, geometric=geometric)
    assert H.levels == cfg.levels
    r = np.random.default_rng(sum(sizes) + colours).standard_normal(H.n)
    want = O.mg_vcycle(cfg, r)
    z = A.Vector(H.n)
    H.vcycle(A.Vector.from_numpy(r), z)
    got = z.numpy()
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("name", ["hpcg_16x16x16_l1_c27", "hpcg_16x16x16_l3_c8", "hpcg_8x12x16_l2_c5",
                                  "hpcg_32x32x32_l3_c8"])
def test_hpcg_matches_reference_fixture(A, golden, name):
    meta = manifest(name)
    g = golden(name)
    nx, ny, nz = meta["sizes"]
    H = A.Hpcg(nx, ny, nz, max_levels=meta["max_levels"], colours=meta["colors"])
    assert H.levels == meta["coarse_levels_built"] + 1
    _, b = rhs_ones(A, nx, ny, nz)
    x = A.Vector(H.n, fill=0.0)
    it, res, _ = H.solve(b, x, meta["max_iter"], meta["tol"], precision=0)
    xh = x.numpy()
    if meta["tol"] > 0:
        assert abs(it - int(g["iterations"][0])) <= 1
        assert res / np.linalg.norm(b.numpy()) <= 1e-10 * 1.5
        assert np.max(np.abs(xh - g["x"])) <= 1e-9
    else:
        assert it == int(g["iterations"][0])
        assert np.max(np.abs(xh - g["x"])) <= 1e-12
    if name == "hpcg_16x16x16_l1_c27":
        # the reference smoke test's bar (tests/smoke/hpcg.cpp:93)
        assert np.linalg.norm(xh - 1.0) < 4e-14


@pytest.mark.parametrize("sizes,levels", [((16, 16, 16), 2), ((24, 32, 16), 2), ((32, 32, 32), 3)])
def test_hpcg_geometric_iterations_match_oracle(A, sizes, levels):
    H = A.Hpcg(*sizes, max_levels=levels, colours=0)
    assert H.proper
    M, b = rhs_ones(A, *sizes)
    x = A.Vector(H.n, fill=0.0)
    it, res, _ = H.solve(b, x, 500, 1e-10)
    cfg = O.hpcg_cfg(*sizes, levels, geometric=True)
    bh = b.numpy()
    xo, ito, reso, _ = O.hpcg(cfg, bh, np.zeros(H.n), 500, 1e-10)
    assert abs(it - ito) <= 1
    assert np.max(np.abs(x.numpy() - xo)) <= 1e-9


def test_hpcg_unpreconditioned_branch(A):
    sizes = (12, 12, 12)
    H = A.Hpcg(*sizes, max_levels=1, colours=0)
    M, b = rhs_ones(A, *sizes)
    x = A.Vector(H.n, fill=0.0)
    it, res, _ = H.solve(b, x, 1000, 1e-10, with_precond=False)
    xo, ito, reso, _ = O.hpcg(O.hpcg_cfg(*sizes, 1, geometric=True), b.numpy(), np.zeros(H.n), 1000, 1e-10,
                              with_precond=False)
    assert abs(it - ito) <= 1
    assert abs(res - reso) <= 1e-3 * reso + 1e-300


def true_relres(A, M, b, x):
    r = A.Vector(M.nrows)
    A.lib().alp_hip_memcpy_d2d(r.ptr, b.ptr, b.nbytes)
    A.lib().alp_hip_scale_f64(M.nrows, -1.0, r.ptr)
    A.mxv(r, M, x)  # r = A x - b
    return np.sqrt(A.dot(r, r)) / np.sqrt(A.dot(b, b))


@pytest.mark.parametrize("precision", [0, 1, 2, 3])
def test_hpcg_mixed_and_fp64_reach_1e10(A, precision):
    """0: FP64; 1: FP32 iterations with reliable FP64 residual updates (the returned
    residual and the stopping test are FP64); 2: FP64 PCG with an FP32 V-cycle;
    3: restarted FP32 solves in FP64 refinement."""
    sizes = (64, 64, 64)
    H = A.Hpcg(*sizes, max_levels=3, colours=0)
    M, b = rhs_ones(A, *sizes)
    x = A.Vector(H.n, fill=0.0)
    it, res, outer = H.solve(b, x, 1000, 1e-10, precision=precision)
    assert true_relres(A, M, b, x) <= 1.5e-10
    assert np.max(np.abs(x.numpy() - 1.0)) < 1e-8
    if precision in (1, 3):
        assert outer >= 2
    if precision == 1:
        # the returned ||r|| is the FP64 residual of the returned x
        assert abs(res - true_relres(A, M, b, x) * np.sqrt(A.dot(b, b))) <= 1e-3 * res
        x0 = A.Vector(H.n, fill=0.0)
        it0, _, _ = H.solve(b, x0, 1000, 1e-10, precision=0)
        assert it <= 1.3 * it0, (it, it0)


def test_hpcg_reliable_updates_stop_at_the_iteration_limit(A):
    sizes = (32, 32, 32)
    H = A.Hpcg(*sizes, max_levels=2, colours=0)
    M, b = rhs_ones(A, *sizes)
    x = A.Vector(H.n, fill=0.0)
    it, res, outer = H.solve(b, x, 5, 1e-10, precision=1)
    assert it == 5
    assert np.isfinite(res) and res > 0
    assert abs(res - true_relres(A, M, b, x) * np.sqrt(A.dot(b, b))) <= 1e-6 * res


def test_hpcg_192_four_levels(A):
    sizes = (192, 192, 192)
    H = A.Hpcg(*sizes, max_levels=3, colours=0)
    assert H.levels == 4 and H.proper and H.nnz == 189_119_224
    M, b = rhs_ones(A, *sizes)
    for precision in (0, 1, 2, 3):
        x = A.Vector(H.n, fill=0.0)
        it, res, outer = H.solve(b, x, 500, 1e-10, precision=precision)
        assert true_relres(A, M, b, x) <= 1.5e-10, precision
        assert np.max(np.abs(x.numpy() - 1.0)) < 1e-7
