"""Writes the golden fixtures in this folder from ALP's own reference backend.

Run in a container that holds the reference sources (``make -C oracle ref``
must have built ``oracle/_ref/libalpref_seq.so``, ALP's sequential
``reference`` backend behind ``oracle/ref_alp.cpp``):

    python tests/golden/make_golden.py              # every fixture
    python tests/golden/make_golden.py transition   # the transition CG group only
    python tests/golden/make_golden.py cg_converge  # CG to 1e-10 at 64^3 and 192^3 only
    python tests/golden/make_golden.py mxm          # grb::mxm cases only
    python tests/golden/make_golden.py mpv          # algorithms::mpv cases only

(the transition group also needs ``oracle/_ref/libalpref_solver.so``).

Every input is generated from a stated seed; outputs are what the reference
returned. The fixtures are data only (inputs where they are not regenerable
from a seed, and expected outputs); ``manifest.json`` records each file's
seed, shape and SHA-256.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as O  # noqa: E402

P = O._ptr
S = ctypes.c_size_t
D = ctypes.c_double


def save(name, manifest, meta, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **arrays)
    with open(path, "rb") as f:
        digest = hashlib.sha256(f.read()).hexdigest()
    manifest[name] = dict(meta, sha256=digest,
                          arrays={k: list(np.asarray(v).shape) for k, v in arrays.items()})


def scaled_stencil(nx, ny, nz, seed):
    """D A D of the 27-point stencil (d ~ U[0.3, 3)) with every row's entries in
    a random order: an SPD CRS whose Jacobi preconditioner matters and whose
    within-row order is not the generator's."""
    A = O.stencil(nx, ny, nz)
    n = len(A.off) - 1
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.3, 3.0, n)
    off = A.off.astype(np.int64)
    idx = A.idx.astype(np.int64)
    rows = np.repeat(np.arange(n), np.diff(off))
    val = d[rows] * A.val * d[idx]
    for i in range(n):
        p = rng.permutation(off[i + 1] - off[i]) + off[i]
        idx[off[i]:off[i + 1]] = idx[p]
        val[off[i]:off[i + 1]] = val[p]
    return off, idx, val, rng.standard_normal(n), rng.random(n)


def transition_cg(manifest):
    """ALP's transition CG C ABI (src/transition/solver.cpp on the reference
    backend, oracle/_ref/libalpref_solver.so): the caller's CRS, x0 and b in,
    x, the return code, iteration count and residual out."""
    R = O.ref_solver()
    Pv = ctypes.c_void_p
    cases = [
        # name, grid, seed, dtype, precond, tol, max_iter, fail_at (precond call that returns 7)
        ("d_plain", (6, 5, 4), 3, np.float64, False, 1e-10, 1000, 0),
        ("d_jacobi", (6, 5, 4), 3, np.float64, True, 1e-10, 1000, 0),
        ("d_maxit", (6, 5, 4), 3, np.float64, False, 1e-10, 10, 0),
        ("d_precond_fails", (6, 5, 4), 3, np.float64, True, 1e-10, 1000, 5),
        ("d_jacobi_large", (12, 10, 9), 8, np.float64, True, 1e-12, 1000, 0),
        ("s_plain", (6, 5, 4), 3, np.float32, False, 1e-5, 1000, 0),
        ("s_jacobi", (6, 5, 4), 3, np.float32, True, 1e-5, 1000, 0),
    ]
    out = {}
    for name, grid, seed, dt, pre, tol, max_iter, fail_at in cases:
        off, idx, val, b, x0 = scaled_stencil(*grid, seed)
        n = len(off) - 1
        sfx = "dii" if dt == np.float64 else "sii"
        ctype = ctypes.c_double if dt == np.float64 else ctypes.c_float
        a, ja, ia = val.astype(dt), idx.astype(np.int32), off.astype(np.int32)
        bb, x = b.astype(dt), x0.astype(dt)
        h = Pv()
        assert getattr(R, "sparse_cg_init_" + sfx)(ctypes.byref(h), ctypes.c_size_t(n), a.ctypes.data_as(Pv),
                                                   ja.ctypes.data_as(Pv), ia.ctypes.data_as(Pv)) == 0
        f = getattr(R, "sparse_cg_set_tolerance_" + sfx)
        f.argtypes = [Pv, ctype]
        f(h, tol)
        f = getattr(R, "sparse_cg_set_max_iter_count_" + sfx)
        f.argtypes = [Pv, ctypes.c_size_t]
        f(h, max_iter)
        cb = None
        if pre:
            diag = np.zeros(n, dt)
            for i in range(n):
                for k in range(ia[i], ia[i + 1]):
                    if ja[k] == i:
                        diag[i] = a[k]
            calls = [0]

            def jacobi(z, r, diag=diag, calls=calls, fail_at=fail_at):
                calls[0] += 1
                if fail_at and calls[0] == fail_at:
                    return 7
                np.divide(r, diag, out=z)
                return 0

            cb = O.precond_callback(jacobi, n, dt)
            f = getattr(R, "sparse_cg_set_preconditioner_" + sfx)
            f.argtypes = [Pv, Pv, Pv]
            assert f(h, ctypes.cast(cb, Pv), None) == 0
        rc = getattr(R, "sparse_cg_solve_" + sfx)(h, x.ctypes.data_as(Pv), bb.ctypes.data_as(Pv))
        it, res = ctypes.c_size_t(), ctype()
        getattr(R, "sparse_cg_get_iter_count_" + sfx)(h, ctypes.byref(it))
        getattr(R, "sparse_cg_get_residual_" + sfx)(h, ctypes.byref(res))
        getattr(R, "sparse_cg_destroy_" + sfx)(h)
        meta = np.array([rc, it.value, max_iter, fail_at, 1 if pre else 0, grid[0], grid[1], grid[2], seed])
        out.update({f"{name}_ia": ia, f"{name}_ja": ja, f"{name}_a": a, f"{name}_b": bb,
                    f"{name}_x0": x0.astype(dt), f"{name}_x": x, f"{name}_meta": meta,
                    f"{name}_tol": np.array([tol], dt), f"{name}_residual": np.array([res.value], dt)})
    save("transition_cg", manifest,
         dict(kind="transition_cg", matrix="D A D of the 27-pt stencil, d ~ U[0.3,3), rows shuffled; "
              "numpy default_rng(seed): d, row permutations, b ~ N(0,1), x0 ~ U[0,1)",
              preconditioner="Jacobi z = r / diag(A) (returns 7 on call fail_at when fail_at > 0)",
              meta="rc, iterations, max_iter, fail_at, preconditioned, nx, ny, nz, seed",
              cases=[c[0] for c in cases]), **out)


def cg_converge(manifest):
    """ALP's conjugate_gradient to a 1e-10 relative residual at the workloads'
    own sizes (x0 = 0, b = A*1, the 27-point stencil): the iteration count, the
    residual and a probe of x (summary only; x is 2 MB at 64^3)."""
    R = O.ref("seq")
    for side in (64, 192):
        n = side ** 3
        x0 = np.zeros(n)
        x = np.zeros(n)
        it, res = S(), D()
        rc = R.ref_stencil_cg(side, side, side, P(x0), None, 1000, 1e-10, P(x), ctypes.byref(it),
                              ctypes.byref(res))
        probe = np.random.default_rng(99).integers(0, n, 64)
        save(f"cg_{side}x{side}x{side}_tol1e-10", manifest,
             dict(kind="cg_summary", sizes=[side] * 3, x0="0", b="A*1", max_iter=1000, tol=1e-10,
                  probe_rng="numpy default_rng(99).integers(0,n,64)"),
             rc=np.array([rc]), iterations=np.array([it.value]), residual=np.array([res.value]),
             x_probe_idx=probe, x_probe=x[probe], x_max_err=np.array([np.max(np.abs(x - 1.0))]))
        print(f"cg {side}^3: rc={rc} iterations={it.value} residual={res.value:.3e}", flush=True)


def random_coo(rows, cols, per_row, rng, values=True):
    """Distinct random coordinates (about per_row per row, some rows empty) in a
    random input order, with N(0,1) values."""
    cells = set()
    for i in range(rows):
        if rng.random() < 0.1:
            continue
        for j in rng.integers(0, cols, int(rng.integers(1, 2 * per_row))):
            cells.add((i, int(j)))
    cells = sorted(cells)
    perm = rng.permutation(len(cells))
    I = np.array([cells[p][0] for p in perm], np.uint32)
    J = np.array([cells[p][1] for p in perm], np.uint32)
    V = rng.standard_normal(len(cells)) if values else np.ones(len(cells))
    return I, J, V


def mxm_cases(manifest):
    """ALP's grb::mxm (reference/blas3.hpp:1744, after a RESIZE call) on matrices
    built by buildMatrixUnique from random coordinates: plus-times, min-plus and
    pattern operands, in-place updates of a non-empty C, transpose_left /
    transpose_right. Inputs (coordinates in input order) and C's CRS and CCS."""
    R = O.ref("seq")
    cases = [
        # name, m, k, n, per_row, seed, ring, transpose, in_place
        ("plus_times", 40, 35, 50, 3, 1, 0, 0, False),
        ("plus_times_inplace", 40, 35, 50, 3, 2, 0, 0, True),
        ("min_plus", 30, 30, 30, 4, 3, 1, 0, False),
        ("min_plus_inplace", 30, 30, 30, 4, 4, 1, 0, True),
        ("pattern", 60, 45, 55, 3, 5, 2, 0, False),
        ("transpose_left", 33, 27, 41, 3, 6, 0, 1, False),
        ("transpose_right", 29, 37, 31, 3, 7, 0, 2, False),
        ("wide", 120, 200, 180, 6, 8, 0, 0, False),
    ]
    out = {}
    for name, m, k, n, per, seed, ring, tr, inplace in cases:
        rng = np.random.default_rng(seed)
        ar, ac = (k, m) if tr == 1 else (m, k)
        br, bc = (n, k) if tr == 2 else (k, n)
        IA, JA, VA = random_coo(ar, ac, per, rng, ring != 2)
        IB, JB, VB = random_coo(br, bc, per, rng, ring != 2)
        if inplace:
            IC, JC, VC = random_coo(m, n, 2, rng)
        else:
            IC, JC, VC = (np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0))
        cap = m * n
        crs_off, ccs_off = np.zeros(m + 1, np.uint64), np.zeros(n + 1, np.uint64)
        crs_idx, ccs_idx = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
        crs_val, ccs_val = np.zeros(cap), np.zeros(cap)
        nz = S()
        rc = R.ref_mxm(m, k, n, len(IA), P(IA), P(JA), P(VA), len(IB), P(IB), P(JB), P(VB), len(IC), P(IC), P(JC),
                       P(VC), ring, tr, P(crs_off), P(crs_idx), P(crs_val), P(ccs_off), P(ccs_idx), P(ccs_val),
                       cap, ctypes.byref(nz))
        assert rc == 0, (name, rc)
        z = nz.value
        meta = np.array([m, k, n, ring, tr, 1 if inplace else 0, seed])
        out.update({f"{name}_meta": meta, f"{name}_IA": IA, f"{name}_JA": JA, f"{name}_VA": VA,
                    f"{name}_IB": IB, f"{name}_JB": JB, f"{name}_VB": VB, f"{name}_IC": IC, f"{name}_JC": JC,
                    f"{name}_VC": VC, f"{name}_crs_off": crs_off, f"{name}_crs_idx": crs_idx[:z],
                    f"{name}_crs_val": crs_val[:z], f"{name}_ccs_off": ccs_off, f"{name}_ccs_idx": ccs_idx[:z],
                    f"{name}_ccs_val": ccs_val[:z]})
        print(f"mxm {name}: {z} entries", flush=True)
    save("mxm", manifest,
         dict(kind="mxm", cases=[c[0] for c in cases],
              meta="m, k, n, ring (0 plus-times f64, 1 min-plus f64, 2 plus-times over void operands), "
                   "transpose (0, 1 transpose_left, 2 transpose_right), in_place, seed",
              inputs="coordinates in input order (buildMatrixUnique), A and B in stored orientation; "
                     "random_coo(rows, cols, per_row, numpy default_rng(seed)) in make_golden.py"),
         **out)


def mpv_cases(manifest):
    """ALP's algorithms::mpv (mpv.hpp:94) over plus-times on a random 60 x 60
    matrix (random_coo, seed 11, about 3 entries per row, some rows empty) and a
    dense v (seed 12), for k = 0 .. 5: u's values and which entries u holds. The
    reference accumulates each product into the vector it overwrites (mxv is
    in place), so for k >= 3 u is not A^k v (k = 3 gives A v + A^3 v)."""
    R = O.ref("seq")
    n = 60
    rng = np.random.default_rng(11)
    I, J, V = random_coo(n, n, 3, rng)
    v = np.random.default_rng(12).standard_normal(n)
    out = dict(I=I, J=J, V=V, v=v)
    for k in range(6):
        u = np.zeros(n)
        a = np.zeros(n, np.uint8)
        rc = R.ref_mpv(n, len(I), P(I), P(J), P(V), k, P(v), P(u), P(a))
        assert rc == 0, (k, rc)
        out[f"u_{k}"] = u
        out[f"assigned_{k}"] = a
    save("mpv", manifest,
         dict(kind="mpv", n=n, ring="plus-times f64", k=list(range(6)),
              inputs="random_coo(60, 60, 3, default_rng(11)) in input order; v = default_rng(12).standard_normal(60)"),
         **out)


def main(only=None):
    manifest = {}
    mpath = os.path.join(HERE, "manifest.json")
    if only:
        with open(mpath) as f:
            manifest = json.load(f)
        {"transition": transition_cg, "cg_converge": cg_converge, "mxm": mxm_cases,
         "mpv": mpv_cases}[only](manifest)
        with open(mpath, "w") as f:
            json.dump(manifest, f, indent=1, sort_keys=True)
        print(f"updated {only}")
        return
    R = O.ref("seq")

    # -- storage: the reference's CRS and CCS of small stencils -------------------
    for (nx, ny, nz) in [(3, 3, 3), (4, 5, 6)]:
        n = nx * ny * nz
        nnz = S()
        R.ref_stencil_nnz(nx, ny, nz, ctypes.byref(nnz))
        out = {}
        for which, tag in ((0, "crs"), (1, "ccs")):
            off = np.zeros(n + 1, np.uint64)
            idx = np.zeros(nnz.value, np.uint32)
            val = np.zeros(nnz.value)
            R.ref_stencil_storage(nx, ny, nz, which, P(off), P(idx), P(val))
            out[f"{tag}_off"], out[f"{tag}_idx"], out[f"{tag}_val"] = off, idx, val
        save(f"stencil_storage_{nx}x{ny}x{nz}", manifest,
             dict(kind="storage", sizes=[nx, ny, nz], diag=26.0, offdiag=-1.0), **out)

    # -- mxv / transposed mxv: y <- y + A x, y0 and x ~ U[0,1) ---------------------
    for (nx, ny, nz, seed) in [(8, 8, 8, 7), (5, 9, 7, 11), (16, 16, 16, 3)]:
        n = nx * ny * nz
        rng = np.random.default_rng(seed)
        x = rng.random(n)
        y0 = rng.random(n)
        y = y0.copy()
        R.ref_stencil_mxv(nx, ny, nz, P(x), P(y))
        yt = y0.copy()
        R.ref_stencil_mxv_transposed(nx, ny, nz, P(x), P(yt))
        save(f"mxv_{nx}x{ny}x{nz}", manifest,
             dict(kind="mxv", sizes=[nx, ny, nz], seed=seed, rng="numpy default_rng: x, then y0"),
             x=x, y0=y0, y=y, yt=yt)

    # -- dot ----------------------------------------------------------------------
    dots = {}
    for n, seed in [(1, 1), (2, 2), (9, 3), (10, 4), (17, 5), (1000, 6), (4097, 7)]:
        rng = np.random.default_rng(seed)
        x = rng.standard_normal(n)
        y = rng.standard_normal(n)
        z = D(0.0)
        R.ref_dot(n, P(x), P(y), ctypes.byref(z))
        dots[f"x_{n}"], dots[f"y_{n}"], dots[f"z_{n}"] = x, y, np.array([z.value])
    save("dot", manifest, dict(kind="dot", rng="numpy default_rng(seed): x, y ~ N(0,1)",
                               cases=[[1, 1], [2, 2], [9, 3], [10, 4], [17, 5], [1000, 6], [4097, 7]]),
         **dots)

    # -- CG: b = A*1, x0 ~ U[0,1) --------------------------------------------------
    for (nx, ny, nz, seed, max_iter, tol) in [(8, 8, 8, 42, 1000, 1e-10), (12, 10, 9, 42, 1000, 1e-10),
                                              (16, 16, 16, 42, 20, 1e-10)]:
        n = nx * ny * nz
        x0 = np.random.default_rng(seed).random(n)
        x = np.zeros(n)
        it, res = S(), D()
        rc = R.ref_stencil_cg(nx, ny, nz, P(x0), None, max_iter, tol, P(x), ctypes.byref(it),
                              ctypes.byref(res))
        save(f"cg_{nx}x{ny}x{nz}_it{max_iter}", manifest,
             dict(kind="cg", sizes=[nx, ny, nz], seed=seed, rng="numpy default_rng(seed).random(n)",
                  b="A*1", max_iter=max_iter, tol=tol),
             x0=x0, x=x, rc=np.array([rc]), iterations=np.array([it.value]),
             residual=np.array([res.value]))

    # -- C1: 64^3 stencil, 50 CG iterations (summary only; x is 2 MB) -------------
    nx = ny = nz = 64
    n = nx * ny * nz
    x0 = np.random.default_rng(42).random(n)
    x = np.zeros(n)
    it, res = S(), D()
    rc = R.ref_stencil_cg(nx, ny, nz, P(x0), None, 50, 1e-300, P(x), ctypes.byref(it),
                          ctypes.byref(res))
    probe = np.random.default_rng(99).integers(0, n, 64)
    save("cg_64x64x64_it50", manifest,
         dict(kind="cg_summary", sizes=[64, 64, 64], seed=42, rng="numpy default_rng(42).random(n)",
              b="A*1", max_iter=50, tol=1e-300, probe_rng="numpy default_rng(99).integers(0,n,64)"),
         rc=np.array([rc]), iterations=np.array([it.value]), residual=np.array([res.value]),
         x_probe_idx=probe, x_probe=x[probe], x_sum=np.array([x.sum()]),
         x_sha256=np.frombuffer(hashlib.sha256(x.tobytes()).digest(), np.uint8))

    # -- HPCG: b = A*1, x0 = 0, mod-k colours -------------------------------------
    for (nx, ny, nz, levels, colors, iters, tol) in [
        (16, 16, 16, 1, 27, 56, 0.0),      # the reference smoke test's default problem
        (16, 16, 16, 3, 8, 12, 0.0),
        (8, 12, 16, 2, 5, 10, 0.0),
        (32, 32, 32, 3, 8, 200, 1e-10),    # to 1e-10 relative residual
    ]:
        n = nx * ny * nz
        x = np.zeros(n)
        it, res = S(), D()
        rc = R.ref_hpcg(nx, ny, nz, levels, colors, 1, 1, iters, tol, 1, P(x), ctypes.byref(it),
                        ctypes.byref(res))
        lv = S()
        R.ref_hpcg_levels(nx, ny, nz, levels, ctypes.byref(lv))
        save(f"hpcg_{nx}x{ny}x{nz}_l{levels}_c{colors}", manifest,
             dict(kind="hpcg", sizes=[nx, ny, nz], max_levels=levels, colors=colors,
                  coarse_levels_built=lv.value, presmooth=1, postsmooth=1, max_iter=iters,
                  tol=tol, b="A*1", x0="0"),
             x=x, rc=np.array([rc]), iterations=np.array([it.value]), residual=np.array([res.value]))

    # -- PageRank and k-hop on graphs -------------------------------------------
    n = 10
    I = np.array(list(range(n)) + [n - 3], np.uint32)
    J = np.array([(i + 1) % n for i in range(n)] + [n - 1], np.uint32)
    pr = np.zeros(n)
    it, q = S(), D()
    rc = R.ref_pagerank(n, len(I), P(I), P(J), 0.85, 1e-7, 1000, P(pr), ctypes.byref(it),
                        ctypes.byref(q))
    save("pagerank_smoke10", manifest,
         dict(kind="pagerank", graph="ring i->(i+1)%10 plus 7->9 (tests/smoke/hook/small_simple_pagerank.cpp:76-82)",
              alpha=0.85, conv=1e-7, max_iter=1000),
         I=I, J=J, pr=pr, rc=np.array([rc]), iterations=np.array([it.value]), quality=np.array([q.value]))

    for scale, seed in [(9, 1), (11, 5)]:
        n = 1 << scale
        s, d = O.rmat_edges(scale, 16 * n, seed=seed)
        I, J = O.symmetrize_dedup(n, s, d)
        perm = np.random.default_rng(seed + 100).permutation(len(I))
        I, J = I[perm], J[perm]
        pr = np.zeros(n)
        it, q = S(), D()
        rc = R.ref_pagerank(n, len(I), P(I), P(J), 0.85, 1e-9, 1000, P(pr), ctypes.byref(it),
                            ctypes.byref(q))
        crs = O.coo_to_compressed(n, n, I, J, which=0)
        src = int(np.argmax(np.diff(crs.off.astype(np.int64))))
        knn = {}
        for k in (1, 2, 3):
            out = np.zeros(n, np.uint8)
            R.ref_knn(n, len(I), P(I), P(J), src, k, P(out))
            knn[f"knn_{k}"] = (out == 1).astype(np.uint8)
        save(f"graph_rmat_s{scale}", manifest,
             dict(kind="graph", generator="orc_rmat_edges scale, 16*2^scale edges, a,b,c=.57,.19,.19",
                  seed=seed, symmetrised=True, dedup=True, self_loops=False,
                  input_order="numpy default_rng(seed+100).permutation", alpha=0.85, conv=1e-9,
                  max_iter=1000, source=src),
             I=I, J=J, pr=pr, rc=np.array([rc]), iterations=np.array([it.value]),
             quality=np.array([q.value]), source=np.array([src]), **knn)

    transition_cg(manifest)
    cg_converge(manifest)
    mxm_cases(manifest)
    mpv_cases(manifest)

    with open(mpath, "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print(f"wrote {len(manifest)} fixtures")


if __name__ == "__main__":
    # `make_golden.py transition` regenerates that group only
    main(sys.argv[1] if len(sys.argv) > 1 else None)
