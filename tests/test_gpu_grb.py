"""The grb:: header layer (alp_amd/include/graphblas.hpp and graphblas/algorithms/)
compiled into an ALP-style C++ program (tests/cpp/grb_program.cpp, built by
`make -C alp_amd`), run on the GPU: buildMatrixUnique, set, mxv, dot,
conjugate_gradient, foldl, knn and a masked lor-land mxv, checked against the
oracle (ALP's reference CG restated, pinned bit-exact to the reference) and
closed forms."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROG = os.path.join(ROOT, "tests", "cpp", "_build", "grb_program")


def closed_form_A1(nx, ny, nz):
    def c(n):
        v = np.full(n, 3.0)
        v[0] = v[-1] = 2.0
        return v
    prod = c(nz)[:, None, None] * c(ny)[None, :, None] * c(nx)[None, None, :]
    return (27.0 - prod).reshape(-1)


@pytest.mark.parametrize("sizes", [(8, 8, 8), (12, 10, 9), (32, 32, 32)])
def test_grb_program(sizes):
    if not os.path.exists(PROG):
        pytest.fail(f"{PROG} was not built (make -C alp_amd)")
    print("running", PROG, sizes, flush=True)
    out = subprocess.run([PROG, *map(str, sizes)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    n = int(np.prod(sizes))
    b = closed_form_A1(*sizes)
    assert res["rc"] == 0 and res["cg_rc"] == 0
    assert res["n"] == n and res["nnz"] == O.stencil(*sizes).nnz
    assert res["nnz_b"] == n
    # b = A 1 is exact in FP64, so b.b is the closed form's exactly up to the dot's order
    assert abs(res["bb"] - float(b @ b)) <= 1e-12 * float(b @ b)
    rc_o, _, it_o, _, _ = O.cg(O.stencil(*sizes), b, np.zeros(n), 1000, 1e-10)
    assert rc_o == 0 and abs(res["iterations"] - it_o) <= 1
    assert res["max_err"] <= 1e-7
    assert res["true_rr"] <= (1e-10 ** 2) * float(b @ b) * 4
    # vertex 0 is a corner: within 2 hops (Chebyshev) lie 3x3x3 vertices, 7 direct neighbours
    assert res["two_hop"] == 27
    assert res["next"] == 7


MTX_PROG = os.path.join(ROOT, "tests", "cpp", "_build", "grb_mtx_program")


def _write_mtx(path, n, pattern, seed):
    rng = np.random.default_rng(seed)
    cells = set()
    while len(cells) < 4 * n:
        i, j = sorted((int(v) for v in rng.integers(0, n, 2)), reverse=True)
        cells.add((i, j))
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate {'pattern' if pattern else 'real'} symmetric\n")
        f.write(f"{n} {n} {len(cells)}\n")
        for i, j in sorted(cells):
            f.write(f"{i + 1} {j + 1}\n" if pattern else f"{i + 1} {j + 1} {rng.standard_normal():.17g}\n")
    return len(cells)


@pytest.mark.parametrize("pattern", [0, 1])
def test_grb_matrix_file_reader_program(tmp_path, pattern):
    """utils::MatrixFileReader + buildMatrixUnique(A, begin, end) in an ALP-style
    program: sizes as the reader reports them, y = A 1 against scipy's reading of
    the same file (row sums; nonzero counts for a pattern file)."""
    import scipy.io
    if not os.path.exists(MTX_PROG):
        pytest.fail(f"{MTX_PROG} was not built (make -C alp_amd)")
    path = os.path.join(str(tmp_path), "m.mtx")
    entries = _write_mtx(path, 150, pattern, 21 + pattern)
    print("running", MTX_PROG, path, pattern, flush=True)
    out = subprocess.run([MTX_PROG, path, str(pattern)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    S = scipy.io.mmread(path).tocsr()
    assert res["rc"] == 0 and res["m"] == 150 and res["n"] == 150
    assert res["entries"] == entries and res["symmetric"] == 1 and res["pattern"] == pattern
    assert res["nz"] == S.nnz and res["nnz"] == S.nnz
    want = np.diff(S.indptr).astype(np.float64) if pattern else np.asarray(S.sum(axis=1)).ravel()
    assert np.max(np.abs(np.asarray(res["y"]) - want)) <= 1e-12 * max(1.0, np.max(np.abs(want)))
