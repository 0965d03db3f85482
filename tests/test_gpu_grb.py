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


SEM_PROG = os.path.join(ROOT, "tests", "cpp", "_build", "grb_semantics_program")
HPCG_PROG = os.path.join(ROOT, "tests", "cpp", "_build", "grb_hpcg_program")


def _run(prog, *args):
    if not os.path.exists(prog):
        pytest.fail(f"{prog} was not built (make -C alp_amd)")
    print("running", prog, *args, flush=True)
    out = subprocess.run([prog, *map(str, args)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


def _vec(res, name):
    return dict(zip((int(i) for i in res[name + "_idx"]), res[name + "_val"]))


def _rows(res, name="A"):
    """The matrix's entries per row, in storage order."""
    rows = {}
    for i, j, v in zip(res[name + "_I"], res[name + "_J"], res[name + "_V"]):
        rows.setdefault(int(i), []).append((int(j), v))
    return rows


def _gather(rows, m, x, mul, add, ident, y0=None, mask=None, add_identity=False):
    """vxm_inner_kernel_gather (reference/blas2.hpp:295-568) row by row, in
    storage order: out = p (+) out; folded into an existing output entry."""
    y = dict(y0 or {})
    for i in range(m):
        if mask is not None and not mask(i):
            continue
        out, hit = ident, False
        if add_identity and i in x:
            out, hit = x[i], True
        for j, a in rows.get(i, []):
            if j in x:
                out, hit = add(mul(a, x[j]), out), True
        if hit:
            y[i] = add(y[i], out) if i in y else out
    return y


def _assert_vec(res, name, want, exact=True):
    got = _vec(res, name)
    assert int(res[name + "_nnz"]) == len(want), (name, res[name + "_nnz"], len(want))
    assert set(got) == set(want), (name, sorted(set(got) ^ set(want)))
    for k, v in want.items():
        if exact:
            assert got[k] == v, (name, k, got[k], v)
        else:
            assert abs(got[k] - v) <= 1e-12 * max(1.0, abs(v)), (name, k, got[k], v)


@pytest.mark.parametrize("seed", [1, 7])
def test_grb_generic_semirings_and_sparse_semantics(seed):
    """The grb:: layer on arbitrary algebras and sparse vectors against brute
    force from the program's own printed inputs (values are small multiples of
    1/4, so every FP64 sum is exact and order-free): mxv / vxm over plus-times,
    min-plus (fresh and accumulated output), max-times, plus-min over int, a
    program-defined operator; lor-land nonzero counts, the pushed sparse input,
    masks and add_identity; OVERLAP, ILLEGAL and MISMATCH; element-wise
    operations, dots and folds under sparsity; a device eWiseLambda; mxm over
    plus-min on int (the generic SpGEMM) and plus-times; masked set; boolean
    vectors built from ranges; a device count over 2^21 entries."""
    r = _run(SEM_PROG, seed)
    assert r["rc"] == 0
    m, n = int(r["m"]), int(r["n"])
    rows = _rows(r)
    ival = {(int(i), int(j)): int(v) for i, j, v in zip(r["I"], r["J"], r["Vi"])}
    irows = {i: [(j, ival[(i, j)]) for j, _ in lst] for i, lst in rows.items()}
    x, xi, xb = _vec(r, "x"), {k: int(v) for k, v in _vec(r, "xi").items()}, _vec(r, "xb")
    inf = float("inf")
    mul, add = (lambda a, b: a * b), (lambda a, b: a + b)
    # 1-3: double semirings
    y1 = _gather(rows, m, x, mul, add, 0.0)
    _assert_vec(r, "y_plus_times", y1)
    _assert_vec(r, "y_min_plus", _gather(rows, m, x, add, min, inf))
    _assert_vec(r, "y_min_plus_acc", _gather(rows, m, x, add, min, inf, y0={i: 1.5 for i in range(m)}))
    _assert_vec(r, "y_max_times", _gather(rows, m, x, mul, max, -inf))
    # 4: plus-min over int
    imax = 2 ** 31 - 1
    _assert_vec(r, "y_plus_min_int", _gather(irows, m, xi, min, add, 0))
    # 5: (add mod 7, times) by vxm: gather over columns
    xm = {k: int(v) for k, v in _vec(r, "xm").items()}
    cols = {}
    for i, lst in irows.items():
        for j, a in lst:
            cols.setdefault(j, []).append((i, a))
    _assert_vec(r, "y_mod7_vxm", _gather(cols, n, xm, lambda a, s: s * a, lambda p, o: (p + o) % 7, 0))
    # 6: lor-land
    lor, land = (lambda a, b: bool(a) or bool(b)), (lambda a, b: bool(a) and bool(b))
    prow = {i: [(j, 1.0) for j, _ in lst] for i, lst in rows.items()}
    _assert_vec(r, "y_lor", _gather(prow, m, xb, land, lor, False))
    _assert_vec(r, "y_lor_push", _gather(prow, m, {3: True}, land, lor, False))
    mask = set(_vec(r, "mask"))
    _assert_vec(r, "y_lor_masked", _gather(prow, m, xb, land, lor, False, mask=lambda i: i not in mask))
    qrow = {i: [(j, 1.0) for j, _ in lst] for i, lst in rows.items() if i < n}
    _assert_vec(r, "y_lor_add_identity", _gather(qrow, n, xb, land, lor, False, add_identity=True))
    # 7: error codes (grb::RC values)
    assert r["rc_overlap"] == 4 and r["rc_dense_sparse"] == 7 and r["rc_mismatch"] == 3
    assert r["rc_safe_overlap_ok"] == 1
    _assert_vec(r, "y_plus_times_twice", {k: 2 * v for k, v in y1.items()})
    # 8: level 1
    u, w = _vec(r, "u"), _vec(r, "w")
    _assert_vec(r, "ewise_op_sub", {k: u[k] - w[k] for k in u if k in w})
    _assert_vec(r, "ewise_monoid_max", {k: max(u.get(k, -inf), w.get(k, -inf)) for k in set(u) | set(w)})
    z3 = dict(u)
    for k in set(x) & set(w):
        z3[k] = z3[k] + x[k] * w[k] if k in z3 else x[k] * w[k]
    _assert_vec(r, "ewise_mul_into_u", z3)
    z4 = dict(u)
    for k in set(x) | set(w):
        t = x[k] + w[k] if (k in x and k in w) else x.get(k, w.get(k))
        z4[k] = z4[k] + t if k in z4 else t
    _assert_vec(r, "ewise_add_into_u", z4)
    _assert_vec(r, "foldl_op_mul", {k: (v * w[k] if k in w else v) for k, v in u.items()})
    both = set(u) & set(w)
    assert r["dot_max_times"] == max([-inf] + [u[k] * w[k] for k in both])
    assert r["dot_plus_times"] == sum(u[k] * w[k] for k in both)
    assert r["fold_sum_int"] == sum(xi.values())
    assert r["fold_max"] == max([-inf] + list(u.values()))
    # 9: eWiseLambda on u's nonzeroes
    _assert_vec(r, "lambda_u", {k: 2 * v + w.get(k, 100.0) for k, v in u.items()})
    # 10: mxm, A times A^T, as maps (the library path's storage order is pinned by
    # the mxm fixture tests)
    def mxm_map(rws, mulf, addf):
        out = {}
        for i, ri in rws.items():
            di = dict(ri)
            for k, rk in rws.items():
                acc = None
                for j, b in rk:
                    if j in di:
                        p = mulf(di[j], b)
                        acc = p if acc is None else addf(acc, p)
                if acc is not None:
                    out[(i, k)] = acc
        return out
    for name, want in (("mxm_plus_min_int", mxm_map(irows, min, add)),
                       ("mxm_plus_times", mxm_map(rows, mul, add))):
        got = {(int(i), int(j)): v for i, j, v in zip(r[name + "_I"], r[name + "_J"], r[name + "_V"])}
        assert got == want, name
    _assert_vec(r, "masked_set", {i: 4.0 for i in mask})
    # 11: the chain with counts left on the device between primitives
    c1 = {k: u[k] - w[k] for k in u if k in w}
    c2 = {k: (c1[k] + u[k] if (k in c1 and k in u) else c1.get(k, u.get(k))) for k in set(c1) | set(u)}
    c3 = dict(c2)
    for k, v in c1.items():
        c3[k] = max(c3[k], v) if k in c3 else v
    c4 = {k: 7.0 for k, v in c1.items() if v != 0}
    for k in set(c3) & set(w):
        t = c3[k] * w[k]
        c4[k] = c4[k] + t if k in c4 else t
    _assert_vec(r, "chain_c1", c1)
    _assert_vec(r, "chain_c2", c2)
    _assert_vec(r, "chain_c3", c3)
    _assert_vec(r, "chain_c4", c4)
    _assert_vec(r, "chain_y", _gather(rows, m, c3, mul, add, 0.0))
    # 12: boolean builds (false entries are entries) and a count over all lanes
    _assert_vec(r, "bool_sparse", {i: float(i % 4 == 0) for i in range(0, n, 2)})
    _assert_vec(r, "bool_dense", {i: float(i % 3 == 0) for i in range(n)})
    L = int(r["long_n"])
    assert int(r["long_nnz"]) == (L + 4) // 5
    # repeated indices fold with subtract in input order; held entries fold first
    _assert_vec(r, "dup_empty", {3: 10.0 - 2 - 4, 5: 1.0 - 8, 0: 0.5})
    _assert_vec(r, "dup_held", {3: 100.0 - 10 - 2 - 4, 5: 1.0 - 8, 0: 0.5, 7: -1.0})
    idx = range((L - 1) % 3, L, 3)
    assert int(r["long_build_nnz"]) == len(idx) and int(r["long_build_bad"]) == 0
    assert r["long_build_sum"] == float(sum(i // 3 for i in idx))


def test_grb_hpcg_program_as_the_reference_smoke_test(golden):
    """build_hpcg_system + hpcg under a grb::Launcher, as tests/smoke/hpcg.cpp
    runs it: 16^3, one coarse level, 27 index-mod colours, 56 iterations; the
    smoke test's bar ||x - 1|| < 4e-14 and the reference's iteration count
    (fixture hpcg_16x16x16_l1_c27). After 56 iterations both residuals are
    round-off (||r||^2 ~ 1e-39 against ||b||^2 ~ 1e4): their values are noise in
    either implementation, so the check is that both sit at that level."""
    g = golden("hpcg_16x16x16_l1_c27")
    r = _run(HPCG_PROG, 16, 16, 16, 1, 27, 56)
    assert r["launcher_rc"] == 0 and r["rc"] == 0 and r["nprocs"] == 1 and r["pid"] == 0
    assert r["iterations"] == int(g["iterations"][0]) == 56
    assert r["residual"] <= 1e-30 and g["residual"][0] <= 1e-30
    assert r["square_norm_diff"] < (4e-14) ** 2
    assert r["pinned_nonzeroes"] == 16 ** 3


PIPE_PROG = os.path.join(ROOT, "tests", "cpp", "_build", "grb_pipeline_program")


@pytest.mark.parametrize("sizes", [(16, 16, 16, 40), (12, 10, 9, 25), (104, 104, 104, 6)])
def test_lazy_primitives_and_pipeline_replay(sizes):
    """Lazy execution (primitives enqueue, the host waits at forcing points) and
    grb::Pipeline (each call's device work captured into hipGraphs, segment by
    segment, and launched by updating the previous call's graphs): every loop of
    tests/cpp/grb_pipeline_program.cpp runs once primitive by primitive and once
    as a Pipeline body from the same start, and the results must be the same
    bits -- a dense Jacobi sweep (one segment per call), scalars that change
    every call, a power iteration with its norm inside the body (the dot splits
    the call into two segments), a value-masked set whose device-side count an
    mxv inside the body asks for, and eight sweeps per call (one 32-launch
    segment). The Jacobi-preconditioned CG in grb::
    primitives must take the oracle CG's iteration count: with D = 26 I the
    preconditioned iterates equal the plain ones in exact arithmetic. The small
    grids run the element-wise primitives fused into one kernel per run of them;
    104^3 (over 2^20 elements) runs each as its own kernel."""
    nx, ny, nz, sweeps = sizes
    r = _run(PIPE_PROG, nx, ny, nz, sweeps)
    assert r["rc"] == 0
    assert r["a_equal"] == 1 and r["b_equal"] == 1 and r["c_equal"] == 1 and r["d_equal"] == 1
    assert r["a_segments"] == 1 and r["c_segments"] == 2 and r["d_segments"] == 2
    # after the first call, each call updates its graph in place
    assert r["a_instantiated"] == 1 and r["a_updated"] == sweeps - 1
    assert r["f_equal"] == 1 and r["f_segments"] == 1
    assert r["b_updated"] == 19
    n = nx * ny * nz
    assert 0 < r["d_nnz"] < n
    b = closed_form_A1(nx, ny, nz)
    rc_o, _, it_o, _, _ = O.cg(O.stencil(nx, ny, nz), b, np.zeros(n), 1000, 1e-10)
    assert r["pcg_rc"] == 0 and rc_o == 0
    assert abs(r["pcg_iterations"] - it_o) <= 1
    assert r["pcg_max_err"] <= 1e-7


ALG_PROG = os.path.join(ROOT, "tests", "cpp", "_build", "grb_algorithms_program")


def _write_general_mtx(path, n, I, J, V=None):
    """MatrixMarket 'coordinate general' in the given entry order (1-based)."""
    with open(path, "w") as f:
        kind = "pattern" if V is None else "real"
        f.write(f"%%MatrixMarket matrix coordinate {kind} general\n{n} {n} {len(I)}\n")
        for t in range(len(I)):
            if V is None:
                f.write(f"{int(I[t]) + 1} {int(J[t]) + 1}\n")
            else:
                f.write(f"{int(I[t]) + 1} {int(J[t]) + 1} {float(V[t]):.17g}\n")


def _hops(n, I, J, source, k, reverse=False):
    """Vertices within k hops of source along the edges i -> j (j -> i reversed):
    knn.hpp's k lor-land products with add_identity, by breadth-first search."""
    adj = [[] for _ in range(n)]
    for i, j in zip(I.tolist(), J.tolist()):
        if reverse:
            adj[j].append(i)
        else:
            adj[i].append(j)
    seen = {source}
    front = [source]
    for _ in range(k):
        nxt = []
        for v in front:
            for w in adj[v]:
                if w not in seen:
                    seen.add(w)
                    nxt.append(w)
        front = nxt
    return sorted(seen)


@pytest.mark.parametrize("name,conv", [("graph_rmat_s9", 1e-9), ("graph_rmat_s11", 1e-9),
                                       ("pagerank_smoke10", 1e-7)])
def test_grb_pagerank_and_knn_device_and_primitive_paths(tmp_path, golden, name, conv):
    """simple_pagerank and knn through the device solvers (no descriptor) and
    through the reference's loops in grb:: primitives (a descriptor that changes
    nothing selects them), on the reference's own fixture graphs read from a
    MatrixMarket file in the fixture's entry order: ranks within 1e-12 of the
    reference's and iterations within one, both ways; k-hop sets equal to the
    reference's (k = 1..3 on the R-MAT graphs), to a breadth-first search on the
    directed ring (knn follows the edges i -> j; transpose_matrix the reverse),
    and MISMATCH for a source out of range (setElement's code in the
    reference)."""
    g = golden(name)
    I, J = g["I"], g["J"]
    n = len(g["pr"])
    src = int(g["source"][0]) if "source" in g.files else 0
    path = os.path.join(str(tmp_path), "g.mtx")
    _write_general_mtx(path, n, I, J)
    r = _run(ALG_PROG, "graph", path, src, conv)
    assert r["rc"] == 0 and r["n"] == n
    it_ref = int(g["iterations"][0])
    for way in ("device", "generic"):
        assert r[f"pr_rc_{way}"] == int(g["rc"][0]) == 0, way
        assert abs(r[f"pr_it_{way}"] - it_ref) <= 1, (way, r[f"pr_it_{way}"], it_ref)
        assert np.max(np.abs(np.array(r[f"pr_{way}"]) - g["pr"])) <= 1e-12, way
    for k in range(4):
        fwd = _hops(n, I, J, src, k)
        rev = _hops(n, I, J, src, k, reverse=True)
        if f"knn_{k}" in g.files:
            assert np.flatnonzero(g[f"knn_{k}"]).tolist() == fwd, k
        assert r[f"knn_device_{k}"] == fwd, k
        assert r[f"knn_generic_{k}"] == fwd, k
        assert r[f"knn_transposed_{k}"] == rev, k
    assert r["knn_rc"] == 0
    assert r["knn_bad_source_rc"] == 3  # MISMATCH


def test_grb_mpv_against_reference_fixture(tmp_path, golden):
    """algorithms::mpv over plus-times for k = 0..5 against ALP's own mpv on the
    same matrix (entries in the same input order) and v (fixture mpv, pinned to
    the loop's closed form in test_oracle.py): which entries u holds, exactly,
    and the values within 1e-14 of the largest."""
    g = golden("mpv")
    n = len(g["v"])
    path = os.path.join(str(tmp_path), "a.mtx")
    _write_general_mtx(path, n, g["I"], g["J"], g["V"])
    vpath = os.path.join(str(tmp_path), "v.txt")
    with open(vpath, "w") as f:
        f.write("\n".join(f"{float(x):.17g}" for x in g["v"]) + "\n")
    r = _run(ALG_PROG, "mpv", path, vpath)
    assert r["rc"] == 0 and r["n"] == n
    for k in range(6):
        assert r[f"rc_{k}"] == 0
        want_a = g[f"assigned_{k}"].astype(bool)
        assert np.array_equal(np.array(r[f"assigned_{k}"]) > 0, want_a), k
        assert r[f"nnz_{k}"] == int(want_a.sum())
        u, want = np.array(r[f"u_{k}"]), g[f"u_{k}"]
        err = np.max(np.abs(u - want)) / np.max(np.abs(want))
        print(f"mpv k={k}: max difference {err:.2e} of the largest", flush=True)
        assert err <= 1e-14, k


REF_ON_HIP = os.path.join(ROOT, "oracle", "_ref", "ref_on_hip")


def _need_ref_on_hip():
    if not os.path.exists(REF_ON_HIP):
        pytest.skip("oracle/_ref/ref_on_hip is built only where the reference sources are (make -C oracle ref)")


@pytest.mark.parametrize("name", ["cg_8x8x8_it1000", "cg_12x10x9_it1000", "cg_16x16x16_it20"])
def test_reference_cg_template_runs_on_the_hip_layer(tmp_path, golden, name):
    """ALP's own conjugate_gradient.hpp, compiled unchanged from the reference
    sources over the hip grb:: layer (oracle/ref_on_hip.cpp): every primitive of
    the reference's loop runs on the device through this layer. Against the
    reference backend's run of the same template (the fixture: b = A 1, x0 from
    the fixture): return code, iterations within one, x within 1e-9 (converged)
    or 1e-10 of the largest entry (20 iterations)."""
    _need_ref_on_hip()
    g = golden(name)
    nx, ny, nz = (int(v) for v in name.split("_")[1].split("x"))
    max_iter = int(name.split("_it")[1])
    x0 = os.path.join(str(tmp_path), "x0.txt")
    with open(x0, "w") as f:
        f.write("\n".join(f"{float(v):.17g}" for v in g["x0"]) + "\n")
    r = _run(REF_ON_HIP, "cg", nx, ny, nz, max_iter, 1e-10, x0)
    assert r["rc"] == 0 and r["cg_rc"] == int(g["rc"][0])
    assert abs(r["iterations"] - int(g["iterations"][0])) <= 1
    x = np.array(r["x"])
    tol = 1e-9 if max_iter == 1000 else 1e-10 * np.max(np.abs(g["x"]))
    assert np.max(np.abs(x - g["x"])) <= tol


def test_reference_knn_and_mpv_templates_run_on_the_hip_layer(tmp_path, golden):
    """ALP's own knn.hpp and mpv.hpp over the hip grb:: layer: the k-hop sets of
    the reference's fixture graphs (k = 1..3 on R-MAT scale 9; the directed ring
    against a breadth-first search) and mpv's u for k = 0..5 against the
    reference backend's (fixture mpv)."""
    _need_ref_on_hip()
    g = golden("graph_rmat_s9")
    n = len(g["pr"])
    path = os.path.join(str(tmp_path), "g.mtx")
    _write_general_mtx(path, n, g["I"], g["J"])
    src = int(g["source"][0])
    for k in (1, 2, 3):
        r = _run(REF_ON_HIP, "knn", path, src, k)
        assert r["rc"] == 0 and r["knn_rc"] == 0
        assert r["reached"] == np.flatnonzero(g[f"knn_{k}"]).tolist(), k
    s = golden("pagerank_smoke10")
    ring = os.path.join(str(tmp_path), "ring.mtx")
    _write_general_mtx(ring, 10, s["I"], s["J"])
    for k in (0, 1, 2, 4):
        r = _run(REF_ON_HIP, "knn", ring, 0, k)
        assert r["reached"] == _hops(10, s["I"], s["J"], 0, k), k
    m = golden("mpv")
    nm = len(m["v"])
    apath = os.path.join(str(tmp_path), "a.mtx")
    _write_general_mtx(apath, nm, m["I"], m["J"], m["V"])
    vpath = os.path.join(str(tmp_path), "v.txt")
    with open(vpath, "w") as f:
        f.write("\n".join(f"{float(x):.17g}" for x in m["v"]) + "\n")
    for k in range(6):
        r = _run(REF_ON_HIP, "mpv", apath, vpath, k)
        assert r["rc"] == 0 and r["mpv_rc"] == 0
        assert np.array_equal(np.array(r["assigned"]) > 0, m[f"assigned_{k}"].astype(bool)), k
        u, want = np.array(r["u"]), m[f"u_{k}"]
        assert np.max(np.abs(u - want)) <= 1e-14 * np.max(np.abs(want)), k
