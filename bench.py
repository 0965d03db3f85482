#!/usr/bin/env python3
"""Benchmark of the hip backend's hot path on one MI355X.

Workload: the 192^3 27-point stencil (7,077,888 rows, 189,119,224 nonzeros;
diagonal 26, off-diagonals -1), stored as device CRS. One step is one
grb::mxv<dense>(y, A, x) under the FP64 plus-times semiring, i.e. y = y + A x
(ALP's accumulate semantics, reference/blas2.hpp:1946).

value = SpMV throughput in GB/s of the algorithmic bytes of the layout the timed
kernel streams (format_bytes_per_launch) over the wall time per step. The default
kernel streams the matrix as index-free diagonal slices:
    8 E (values, E padded slot entries) + 393 S (per slice: a 4 B presence mask per
    row, 32 slot offsets, a flag, the entry offset) + 8 n (x) + 16 n (y r+w)
  = 1,731,557,376 B at 192^3.
roofline.achieved / roofline.frac are the same bytes over the kernel's HIP-event
time per launch, so frac <= 1 is a fraction of the 8 TB/s HBM peak. The metric's
CRS algorithmic bytes (12 nnz + 8 (n+1) + 8 n + 16 n = 2,495,923,112 B) over the
same time are reported separately as crs_equivalent_gbs (they exceed what the
hardware moves). crs_path times the CSR-stream kernel (kernel 2) forced on the same
matrix, in CRS bytes. roofline.traffic is the measured HBM traffic per launch. The
solve legs (FP64 and mixed-precision CG / HPCG to 1e-10 relative residual) are
reported beside it (`solves`), and `configs` holds the other workloads: 50 CG
iterations at 64^3, BFS / PageRank / k-hop on an R-MAT graph of scale 22, 64
batched systems x 8 right-hand sides, and the differentiable solve at 128^3.
`cpu_baseline` times ALP's own reference_omp on the host: the same SpMV, and
(`cpu_baseline.solves`) bounded samples of its CG and HPCG. Inputs are resident in
HBM before the timed region.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
    python bench.py --profile-solve {cg,hpcg,batch} --precision P --iters N   (a short
        solve for kernel traces)
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "CG solve time & SpMV GB/s (% HBM peak): mixed vs FP64 at 1e-10 rel residual"
HBM_PEAK_GBS = 8000.0
KERNEL_NAMES = {1: "csr-vector", 2: "csr-stream", 3: "csr-wave", 4: "sliced-ell", 5: "sliced-ell-diagonal"}


def spmv_bytes(n: int, nnz: int, accumulate: bool = True) -> int:
    return 12 * nnz + 8 * (n + 1) + 8 * n + (16 if accumulate else 8) * n


def format_bytes(n: int, nnz: int, kernel: int, layout, accumulate: bool = True) -> int:
    """Bytes one FP64 SpMV launch must move in the layout the kernel streams.
    Diagonal slices: values (8 B per padded entry), a 4 B presence mask per row,
    per slice its flag, 32 slot offsets and its entry offset; sliced-ELL slices add
    4 B column indices and read the row offsets; x once and y as for CRS."""
    slices, entries, dia = layout
    vec = 8 * n + (16 if accumulate else 8) * n
    if kernel in (4, 5) and slices:
        share = dia / slices if kernel == 5 else 0.0
        per_slice = 8 + (share * (1 + 32 * 4 + 64 * 4))
        sell_part = (1.0 - share) * (4 * entries + 8 * (n + 1))
        return int(8 * entries + sell_part + per_slice * slices + vec)
    return spmv_bytes(n, nnz, accumulate)


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(side: int, nnz: int, n: int) -> dict:
    """ALP's own reference_omp (oracle/_ref) timed on this host; else the oracle port."""
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if O.ref_available("omp"):
        R = O.ref("omp")
        reps = 10
        ms = R.ref_time_stencil_mxv(side, side, side, reps)
        return {"value": round(spmv_bytes(n, nnz) / (ms * 1e-3) / 1e9, 3), "unit": "GB/s",
                "cores": threads, "kind": "reference",
                "sample": f"ALP reference_omp grb::mxv<dense> on the {side}^3 stencil, median of "
                          f"{reps} calls after 1 warm-up ({ms:.2f} ms per call)"}
    A = O.stencil(side, side, side)
    x = np.ones(A.nrows)
    reps = 3
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.spmv(A, x)
        t.append(time.perf_counter() - t0)
    ms = 1e3 * sorted(t)[reps // 2]
    return {"value": round(spmv_bytes(n, nnz) / (ms * 1e-3) / 1e9, 3), "unit": "GB/s", "cores": 1,
            "kind": "port", "sample": f"oracle orc_spmv_f64 on the {side}^3 stencil, median of {reps}"}


def cpu_solve_reference(side: int, cg_iterations: int, hpcg_iterations: int):
    """The reference's own solvers (oracle/_ref, reference_omp) on this host: the
    whole FP64 CG solve to 1e-10 at side^3 (x0 = 0, b = A 1, system build not
    timed; about 5 s on 16 threads), so its iteration count checks the device's
    and its time is measured, not estimated; and a bounded HPCG sample (2
    iterations, 4 levels, 8 colours; a whole solve would take about 40 s) scaled
    by the device's FP64 iteration count."""
    import oracle as O
    if not O.ref_available("omp"):
        return None
    R = O.ref("omp")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    it = ctypes.c_size_t()
    out = {"kind": "reference", "cores": threads}
    ms = R.ref_time_stencil_cg(side, side, side, 1000, 1e-10, ctypes.byref(it))
    if ms > 0 and it.value:
        out["cg_fp64"] = {"ms": round(ms, 1), "iterations": it.value, "device_iterations": cg_iterations,
                          "ms_per_iteration": round(ms / it.value, 2), "timed": "whole solve to 1e-10"}
    ms = R.ref_time_hpcg(side, side, side, 3, 8, 2, ctypes.byref(it))
    if ms > 0 and it.value:
        per = ms / it.value
        out["hpcg_fp64"] = {"iterations_timed": it.value, "ms_per_iteration": round(per, 2),
                            "est_ms_full_solve": round(per * hpcg_iterations, 1), "iterations": hpcg_iterations}
    out["sample"] = (f"ALP reference_omp conjugate_gradient to 1e-10 (whole solve, timed) and hpcg (2 iterations, "
                     f"4 levels, 8 colours; est_ms_full_solve = ms_per_iteration x the device's FP64 iteration "
                     f"count) on the {side}^3 stencil")
    return out


def measured_traffic(kernel: int, lanes: int, rows: int):
    """HBM bytes per launch of this SpMV configuration from the committed rocprofv3
    FETCH_SIZE + WRITE_SIZE summaries (profiles/*_hbm_counters.md, calibrated on a
    copy kernel of known size), or None when none covers it."""
    import glob
    import re
    pat = re.compile(r"traffic_bytes_per_launch:\s*(\d+)\s*\(kernel=(\d+) lanes=(\d+) rows=(\d+), accumulate\)")
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_hbm_counters.md"))):
        with open(path) as f:
            for m in pat.finditer(f.read()):
                if (int(m.group(2)), int(m.group(3)), int(m.group(4))) == (kernel, lanes, rows):
                    return int(m.group(1)), os.path.relpath(path, ROOT)
    return None


def timed(fn, sync):
    sync()
    t0 = time.perf_counter()
    out = fn()
    sync()
    return out, time.perf_counter() - t0


def leg_grb_loops() -> dict:
    """A program's own loops in grb:: primitives (tests/cpp/grb_pipeline_program,
    built by `make -C alp_amd`): a Jacobi sweep of four primitives and a
    Jacobi-preconditioned CG iteration at 64^3, with lazy execution and fused
    element-wise runs; the same sweep as a grb::Pipeline of hipGraphs."""
    import subprocess
    prog = os.path.join(ROOT, "tests", "cpp", "_build", "grb_pipeline_program")
    if not os.path.exists(prog):
        return {"skipped": "tests/cpp/_build/grb_pipeline_program not built"}
    out = subprocess.run([prog, "64", "64", "64", "48"], capture_output=True, text=True, timeout=300)
    if out.returncode != 0:
        raise RuntimeError("grb_pipeline_program failed: " + out.stdout[-500:] + out.stderr[-500:])
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"grid": "64^3", "jacobi_sweep_us": r["a_eager_us"], "jacobi_sweep_pipeline_us": r["a_pipeline_us"],
            "pcg_iterations": r["pcg_iterations"], "pcg_us_per_iteration": r["pcg_us_per_iteration"],
            "bit_identical_pipeline": bool(r["a_equal"] and r["b_equal"] and r["c_equal"] and r["d_equal"])}


def leg_c1(alp, L) -> dict:
    """64^3 stencil, b = A 1, x0 ~ U[0,1) (numpy seed 42), 50 unpreconditioned
    FP64 CG iterations; GB/s counts the SpMV overwrite form plus 72 B/row of
    fused BLAS1 (update 48, direction 24) per iteration. The CPU reference runs
    the same 50 iterations (from x0 = 0; the count, not x0, sets the work)."""
    side = 64
    A = alp.Matrix.stencil27(side, side, side)
    n, nnz = A.nrows, A.nnz
    b = alp.Vector(n, fill=0.0)
    alp.mxv(b, A, alp.Vector(n, fill=1.0))
    x0 = np.random.default_rng(42).random(n)
    times = []
    for _ in range(5):
        x = alp.Vector.from_numpy(x0)
        (rc, it, res), dt = timed(lambda: alp.cg(A, b, x, 50, 1e-300), alp.sync)
        times.append(dt)
    ms = 1e3 * sorted(times)[2]
    per_it = spmv_bytes(n, nnz, accumulate=False) + 72 * n
    out = {"ms": round(ms, 3), "iterations": it, "residual": res,
           "gbs_per_iteration": round(per_it * it / (ms * 1e-3) / 1e9, 1)}
    import oracle as O
    if O.ref_available("omp"):
        R = O.ref("omp")
        it_ref = ctypes.c_size_t()
        ref_ms = R.ref_time_stencil_cg(side, side, side, 50, 1e-300, ctypes.byref(it_ref))
        out["cpu_reference_omp_ms"] = round(ref_ms, 2)
        out["speedup_vs_cpu"] = round(ref_ms / ms, 1)
    return out


def leg_c2(alp, scale: int = 22) -> dict:
    """R-MAT scale 22, edge factor 16, (a, b, c) = (.57, .19, .19), seed 1, symmetrised,
    deduplicated, no self-loops; BFS from the highest-degree vertex; PageRank
    (0.85, 1e-9); knn (k-hop reachability) 1 and 3 hops from the same vertex."""
    n = 1 << scale
    (src, dst), t_gen = timed(lambda: alp.rmat_edges(scale, 16 * n, seed=1), alp.sync)
    G, t_build = timed(lambda: alp.graph_from_edges(n, src, dst), alp.sync)
    del src, dst
    nnz = G.nnz
    s, deg = alp.max_degree_vertex(G)
    # the levels go to a caller-owned vector, as grb:: algorithms take theirs
    lev = alp.Vector(n, np.int32)
    alp.bfs(G, s, out=lev)  # warm-up
    times = []
    for _ in range(5):
        (lev, nl, reached), dt = timed(lambda: alp.bfs(G, s, out=lev), alp.sync)
        times.append(dt)
    bfs_ms = 1e3 * sorted(times)[2]
    # warm-up: the first call builds the gather pattern in out-degree order, once
    # per graph (as the reference builds its CCS with the matrix)
    _, t_pr_first = timed(lambda: alp.pagerank(G, 0.85, 1e-9, 1000), alp.sync)
    (rc, pr, it, q), t_pr = timed(lambda: alp.pagerank(G, 0.85, 1e-9, 1000), alp.sync)
    # the default product is the row-segment CSR-stream variant; the same call
    # with the storage-order variant, for comparison
    G.set_kernel(2, 8, 8)
    alp.pagerank(G, 0.85, 1e-9, 1000)
    (rc2, pr2, it2, _), t_pr2 = timed(lambda: alp.pagerank(G, 0.85, 1e-9, 1000), alp.sync)
    G.set_kernel(0, 0, 0)
    order_diff = float(np.max(np.abs(pr2.numpy() - pr.numpy())))
    pr_bytes = 4 * nnz + 8 * (n + 1) + 32 * n
    # knn (k-hop reachability) from the same source, 1 and 3 hops
    kh = {}
    mask = alp.Vector(n, np.uint8)
    for k in (1, 3):
        alp.khop(G, s, k, out=mask)  # warm-up
        out, t_kh = timed(lambda: alp.khop(G, s, k, out=mask), alp.sync)
        kh[k] = (round(1e3 * t_kh, 3), int(np.count_nonzero(out.numpy())))
    return {"n": n, "nnz": nnz, "source": s, "source_degree": deg, "gen_ms": round(1e3 * t_gen, 2),
            "build_ms": round(1e3 * t_build, 2), "bfs_ms": round(bfs_ms, 3), "bfs_levels": nl,
            "bfs_reached": reached, "bfs_gteps": round(nnz / 2 / (bfs_ms * 1e-3) / 1e9, 2),
            "pagerank_rc": rc, "pagerank_iterations": it, "pagerank_ms": round(1e3 * t_pr, 2),
            "pagerank_gbs": round(pr_bytes * it / t_pr / 1e9, 1), "pagerank_first_call_ms": round(1e3 * t_pr_first, 2),
            "pagerank_storage_order_ms": round(1e3 * t_pr2, 2), "pagerank_storage_order_rc_iterations": [rc2, it2],
            "pagerank_storage_order_max_abs_diff": order_diff,
            "khop1_ms": kh[1][0], "khop1_reached": kh[1][1], "khop3_ms": kh[3][0], "khop3_reached": kh[3][1]}


def leg_c4(alp) -> dict:
    """64 anisotropic 7-point Laplacians, sides U{16..96} per axis, coefficients
    lognormal(0, 0.5) (numpy seed 2024), 8 N(0,1) right-hand sides each, to 1e-8."""
    rng = np.random.default_rng(2024)
    nx, ny, nz = (rng.integers(16, 97, 64) for _ in range(3))
    cx, cy, cz = (rng.lognormal(0.0, 0.5, 64) for _ in range(3))
    B = alp.Batch(nx, ny, nz, cx, cy, cz, 8)
    w = B.autotune()
    rhs = alp.Vector.from_numpy(rng.standard_normal(B.nrows * 8))
    out = {"systems": 64, "rhs": 8, "rows": B.nrows, "nnz": B.nnz, "spmm_lanes": w}
    for precision, tag in ((0, "fp64"), (1, "mixed"), (2, "mixed_restarted")):
        X = alp.Vector(B.nrows * 8, fill=0.0)
        B.solve(rhs, X, 1e-8, 20000, precision)  # warm-up (graph capture)
        runs = [timed(lambda: B.solve(rhs, X, 1e-8, 20000, precision), alp.sync) for _ in range(3)]
        (rc, its, rel, outer), dt = sorted(runs, key=lambda r: r[1])[1]  # the median of 3
        out[tag] = {"rc": rc, "ms": round(1e3 * dt, 2), "max_iterations": int(its.max()),
                    "max_rel_residual": float(rel.max()), "refinements": outer}
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["ms"] / out["mixed"]["ms"], 3)
    B.free()
    return out


def profile_solve(alp, args) -> int:
    """One solve of exactly args.iters iterations (tolerance 1e-30) at
    args.precision, for kernel traces: the per-kernel times of an iteration
    without the timing loops of the full run."""
    what, prec, its = args.profile_solve, args.precision, args.iters
    if what == "adjoint":
        side = 128
        n = side ** 3
        rng = np.random.default_rng(5)
        theta = alp.Vector.from_numpy(rng.lognormal(0.0, 1.0, 13 * n))
        b = alp.Vector.from_numpy(rng.standard_normal(n))
        A = alp.Matrix.conductance27(side, side, side, theta, 1e-2,
                                     alp.MAT_WITH_CCS | (alp.MAT_WITH_F32 if prec else 0))
        x, lam, dth = alp.Vector(n, fill=0.0), alp.Vector(n, fill=0.0), alp.Vector(13 * n)
        rc1, it1, _ = alp.cg(A, b, x, its, 1e-30, prec)
        rc2, it2, _ = alp.cg(A, b, lam, its, 1e-30, prec, transpose=True)
        alp.check(alp.lib().alp_hip_grad_conductance27(side, side, side, lam.ptr, x.ptr, dth.ptr), "grad")
        alp.sync()
        log(f"[bench] profile-solve adjoint precision={prec} forward_iterations={it1} adjoint_iterations={it2}")
        return 0
    if what == "graph":
        n = 1 << 22
        src, dst = alp.rmat_edges(22, 16 * n, seed=1)
        G = alp.graph_from_edges(n, src, dst)
        del src, dst
        s, _ = alp.max_degree_vertex(G)
        alp.pagerank(G, 0.85, 1e-9, 1000)  # builds the CCS
        lev, nl, reached = alp.bfs(G, s)
        rc, pr, it, q = alp.pagerank(G, 0.85, 1e-9, 1000)
        alp.khop(G, s, 3)
        alp.sync()
        log(f"[bench] profile-solve graph bfs_levels={nl} pagerank_iterations={it} rc={rc}")
        return 0
    if what == "batch":
        rng = np.random.default_rng(2024)
        nx, ny, nz = (rng.integers(16, 97, 64) for _ in range(3))
        cx, cy, cz = (rng.lognormal(0.0, 0.5, 64) for _ in range(3))
        B = alp.Batch(nx, ny, nz, cx, cy, cz, 8)
        w = B.autotune()
        rhs = alp.Vector.from_numpy(rng.standard_normal(B.nrows * 8))
        X = alp.Vector(B.nrows * 8, fill=0.0)
        rc = B.solve(rhs, X, 1e-30, its, prec)[0]
        alp.sync()
        log(f"[bench] profile-solve batch precision={prec} iterations={its} lanes={w} rc={rc}")
        return 0
    side = args.side
    A = alp.Matrix.stencil27(side, side, side, flags=alp.MAT_WITH_F32)
    b = alp.Vector(A.nrows, fill=0.0)
    alp.check(alp.mxv(b, A, alp.Vector(A.nrows, fill=1.0)), "mxv")
    x = alp.Vector(A.nrows, fill=0.0)
    if what == "cg":
        rc, it, _ = alp.cg(A, b, x, its, 1e-30, prec)
    else:
        H = alp.Hpcg(side, side, side, max_levels=3, colours=0)
        it, _, _ = H.solve(b, x, its, 1e-30, prec)
        rc = 0
    alp.sync()
    log(f"[bench] profile-solve {what} {side}^3 precision={prec} iterations={it} rc={rc}")
    return 0


def leg_c5(alp, L, side: int = 128) -> dict:
    """Differentiable solve on the 128^3 27-point graph: theta ~ lognormal(0, 1)
    per edge (numpy seed 5), A(theta) = weighted Laplacian + 1e-2 I, b ~ N(0,1),
    x* from a second theta (seed 6); loss ||x - x*||^2. Timed: forward CG, the
    adjoint (transposed) solve and the dtheta scatter, to 1e-10."""
    n = side ** 3
    rng = np.random.default_rng(5)
    theta = alp.Vector.from_numpy(rng.lognormal(0.0, 1.0, 13 * n))
    theta_ref = alp.Vector.from_numpy(np.random.default_rng(6).lognormal(0.0, 1.0, 13 * n))
    b = alp.Vector.from_numpy(rng.standard_normal(n))
    Aref = alp.Matrix.conductance27(side, side, side, theta_ref, 1e-2)
    xs = alp.Vector(n, fill=0.0)
    alp.cg(Aref, b, xs, 100000, 1e-12)
    Aref.free()
    xstar = xs.numpy()
    out = {"n": n, "edges_slots": 13 * n}
    grads = {}
    for precision, tag in ((0, "fp64"), (1, "mixed")):
        A = alp.Matrix.conductance27(side, side, side, theta, 1e-2,
                                     alp.MAT_WITH_CCS | (alp.MAT_WITH_F32 if precision else 0))
        x = alp.Vector(n, fill=0.0)
        lam = alp.Vector(n, fill=0.0)
        dth = alp.Vector(13 * n)
        alp.sync()
        t0 = time.perf_counter()
        rc1, it1, _ = alp.cg(A, b, x, 100000, 1e-10, precision)
        t1 = time.perf_counter()
        g = alp.Vector.from_numpy(2.0 * (x.numpy() - xstar))  # dL/dx (host, untimed below)
        t2 = time.perf_counter()
        rc2, it2, _ = alp.cg(A, g, lam, 100000, 1e-10, precision, transpose=True)
        alp.check(L.alp_hip_grad_conductance27(side, side, side, lam.ptr, x.ptr, dth.ptr), "grad")
        alp.sync()
        t3 = time.perf_counter()
        grads[tag] = dth.numpy()
        out[tag] = {"rc": [rc1, rc2], "forward_iterations": it1, "adjoint_iterations": it2,
                    "forward_ms": round(1e3 * (t1 - t0), 2), "adjoint_and_scatter_ms": round(1e3 * (t3 - t2), 2)}
        A.free()
    gm = np.max(np.abs(grads["fp64"]))
    out["grad_rel_diff_mixed_vs_fp64"] = float(np.max(np.abs(grads["mixed"] - grads["fp64"])) / gm)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--side", type=int, default=192)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the other workloads (64^3 CG, R-MAT graphs, batched CG, adjoint)")
    ap.add_argument("--skip-restarted", action="store_true",
                    help="leave out the restarted-refinement solve modes (CG precision 2, HPCG precision 3)")
    ap.add_argument("--autotune", action="store_true",
                    help="time every SpMV kernel variant on the matrix and keep the fastest "
                         "(default: the library's own choice)")
    ap.add_argument("--kernel", type=int, default=0,
                    help="skip the autotuner: 1 = CSR-vector, 2 = CSR-stream (with --lanes/--rows)")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a plain streaming kernel of known size (f64 -> f32 conversion of "
                         "2^28 elements: 2 GiB read, 1 GiB written) for HBM counter calibration")
    ap.add_argument("--profile-solve", choices=("cg", "hpcg", "batch", "graph", "adjoint"), default=None,
                    help="run one short solve of --iters iterations (tolerance 1e-30, so every "
                         "iteration runs) at --precision and exit: a workload for kernel traces "
                         "(cg, hpcg: the 192^3 stencil; batch: the 64-system batch; graph: BFS, "
                         "PageRank and 3-hop k-hop on the scale-22 R-MAT graph; adjoint: forward and transposed "
                         "solves on the 128^3 conductance matrix, then the dtheta scatter)")
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--iters", type=int, default=48)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path; run with --gpus 1",
              file=sys.stderr)
        return 2

    import alp_amd as alp
    alp.init(0)
    L = alp.lib()
    side = args.side
    if args.profile_solve:
        return profile_solve(alp, args)
    log(f"[bench] building {side}^3 27-point stencil on the device")
    alp.sync()
    t_build0 = time.perf_counter()
    A = alp.Matrix.stencil27(side, side, side, flags=alp.MAT_WITH_F32)
    alp.sync()
    build_ms = 1e3 * (time.perf_counter() - t_build0)
    n, nnz = A.nrows, A.nnz
    if args.kernel:
        A.set_kernel(args.kernel, args.lanes, args.rows)
        kernel, lanes, rows = A.get_kernel()
    elif args.autotune:
        kernel, lanes, rows = A.autotune()
    else:
        kernel, lanes, rows = A.get_kernel()
    layout = A.layout(0)
    log(f"[bench] n={n} nnz={nnz} kernel={KERNEL_NAMES.get(kernel, kernel)} lanes={lanes} rows={rows}")
    x = alp.Vector.from_numpy(np.random.default_rng(0).random(n))
    y = alp.Vector(n, fill=0.0)

    for _ in range(args.warmup):
        alp.check(alp.mxv(y, A, x, descr=alp.DENSE), "mxv")
    alp.sync()
    ev = ctypes.c_double()
    t0 = time.perf_counter()
    alp.check(L.alp_hip_timer_start(), "timer_start")
    for _ in range(args.steps):
        L.alp_hip_mxv(A.h, alp.PLUS_TIMES_F64, alp.DENSE, y.ptr, x.ptr, None)
    alp.check(L.alp_hip_timer_stop(ctypes.byref(ev)), "timer_stop")
    alp.sync()
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    kern_ms = ev.value / args.steps
    # 20 launches timed one by one (HIP events around each), for the spread
    single = []
    for _ in range(20):
        alp.check(L.alp_hip_timer_start(), "timer_start")
        L.alp_hip_mxv(A.h, alp.PLUS_TIMES_F64, alp.DENSE, y.ptr, x.ptr, None)
        alp.check(L.alp_hip_timer_stop(ctypes.byref(ev)), "timer_stop")
        single.append(ev.value * 1e3)
    single.sort()
    per_launch = {"min": round(single[0], 2), "median": round(0.5 * (single[9] + single[10]), 2),
                  "max": round(single[-1], 2), "launches": 20, "unit": "us",
                  "note": "each launch between its own HIP events, after the timed region"}
    crs_bytes = spmv_bytes(n, nnz)
    fbytes = format_bytes(n, nnz, kernel, layout)
    achieved = fbytes / (kern_ms * 1e-3) / 1e9
    crs_equiv = crs_bytes / (kern_ms * 1e-3) / 1e9
    value = fbytes / (ms_step * 1e-3) / 1e9
    log(f"[bench] mxv: {kern_ms * 1e3:.1f} us/launch (events), {ms_step * 1e3:.1f} us/step (wall), "
        f"{achieved:.0f} GB/s of the streamed layout ({fbytes} B; slices={layout[0]} entries={layout[1]} "
        f"diagonal={layout[2]}), {crs_equiv:.0f} GB/s CRS-equivalent")

    # the CRS kernel any non-stencil matrix gets (CSR-stream), forced on the same matrix
    crs_path = None
    if kernel in (4, 5) and not args.kernel:
        # the CRS kernel any matrix without a sliced form gets, forced on the same
        # matrix: CSR-stream in storage order (variant 8) and with row segments
        # staged in LDS (variant 64); the faster one is crs_path
        variants = {}
        for name, crs_rows in (("csr-stream", 8), ("csr-stream-rowseg", 64)):
            A.set_kernel(2, 0, crs_rows)
            for _ in range(max(2, args.warmup // 4)):
                L.alp_hip_mxv(A.h, alp.PLUS_TIMES_F64, alp.DENSE, y.ptr, x.ptr, None)
            alp.check(L.alp_hip_timer_start(), "timer_start")
            crs_steps = max(10, args.steps // 4)
            for _ in range(crs_steps):
                L.alp_hip_mxv(A.h, alp.PLUS_TIMES_F64, alp.DENSE, y.ptr, x.ptr, None)
            alp.check(L.alp_hip_timer_stop(ctypes.byref(ev)), "timer_stop")
            c_ms = ev.value / crs_steps
            c_gbs = crs_bytes / (c_ms * 1e-3) / 1e9
            variants[name] = {"kernel": name, "rows": crs_rows, "launches": crs_steps,
                              "us_per_launch": round(c_ms * 1e3, 2), "bytes_per_launch": crs_bytes,
                              "achieved_gbs": round(c_gbs, 1), "frac": round(c_gbs / HBM_PEAK_GBS, 4)}
        crs_path = dict(min(variants.values(), key=lambda v: v["us_per_launch"]))
        crs_path["variants_us_per_launch"] = {k: v["us_per_launch"] for k, v in variants.items()}
        tr2 = measured_traffic(2, 8, crs_path["rows"])
        if tr2 is not None:
            crs_path["traffic"], crs_path["traffic_source"] = tr2
        A.set_kernel(kernel, lanes, rows)
        log(f"[bench] mxv CRS kernels forced: {crs_path['variants_us_per_launch']} us/launch; "
            f"{crs_path['kernel']} {crs_path['achieved_gbs']:.0f} GB/s")

    # the overwrite form (set(y,0) folded into the product: y write-only)
    alp.sync()
    alp.check(L.alp_hip_timer_start(), "timer_start")
    for _ in range(args.steps):
        L.alp_hip_mxv(A.h, alp.PLUS_TIMES_F64, alp.DENSE | alp.OVERWRITE, y.ptr, x.ptr, None)
    alp.check(L.alp_hip_timer_stop(ctypes.byref(ev)), "timer_stop")
    ow_ms = ev.value / args.steps
    ow_gbs = spmv_bytes(n, nnz, accumulate=False) / (ow_ms * 1e-3) / 1e9

    if args.calibrate:
        m = 1 << 28
        src = alp.Vector(m, fill=1.0)
        dst = alp.Vector(m, np.float32)
        for _ in range(3):
            L.alp_hip_convert_f64_to_f32(m, src.ptr, dst.ptr)
        alp.sync()
        del src, dst

    solves = {}
    if not args.no_solve:
        b = alp.Vector(n, fill=0.0)
        alp.mxv(b, A, alp.Vector(n, fill=1.0))
        cg_modes = ((0, "fp64"), (1, "mixed"), (3, "mixed_groupwise")) + \
            (() if args.skip_restarted else ((2, "mixed_restarted"),))
        for precision, tag in cg_modes:
            xs = alp.Vector(n, fill=0.0)
            alp.cg(A, b, xs, 2, 1e-10, precision)  # warm-up (work vectors, graph capture)
            times = []
            for rep in range(3):
                xs.set(0.0)
                alp.sync()
                t0 = time.perf_counter()
                rc, it, res = alp.cg(A, b, xs, 20000, 1e-10, precision)
                times.append(time.perf_counter() - t0)
            dt = sorted(times)[1]
            err = float(np.max(np.abs(xs.numpy() - 1.0)))
            solves[f"cg_{tag}"] = {"ms": round(dt * 1e3, 2), "iterations": it, "rc": rc,
                                   "rel_residual": float(np.sqrt(res) / np.sqrt(float(alp.dot(b, b)))),
                                   "max_err_vs_exact": err}
            log(f"[bench] CG {tag}: rc={rc} it={it} {dt * 1e3:.1f} ms err={err:.2e}")
        if "cg_fp64" in solves and "cg_mixed" in solves:
            solves["cg_speedup_mixed_vs_fp64"] = round(solves["cg_fp64"]["ms"] / solves["cg_mixed"]["ms"], 3)
            solves["cg_mixed_scheme"] = ("FP32 iterations (FP32 A, r, p, Ap; FP64 row sums) with the solution update "
                                         "x += alpha p in FP64 every iteration and reliable FP64 residual updates; "
                                         "mixed_groupwise = the same with a segment's increments of x accumulated "
                                         "in FP32 and added to the FP64 x before each FP64 residual; "
                                         "mixed_restarted = FP32 inner solves in FP64 iterative refinement")
            if "cg_mixed_groupwise" in solves:
                solves["cg_speedup_mixed_groupwise_vs_fp64"] = round(
                    solves["cg_fp64"]["ms"] / solves["cg_mixed_groupwise"]["ms"], 3)
        # HPCG: 4 grids (192/96/48/24), 8-colour symmetric Gauss-Seidel, b = A*1, x0 = 0
        log("[bench] building the HPCG hierarchy")
        H = alp.Hpcg(side, side, side, max_levels=3, colours=0)
        bnorm = np.sqrt(float(alp.dot(b, b)))
        hpcg_modes = ((0, "fp64"), (1, "mixed"), (2, "mixed_pc")) + (() if args.skip_restarted else ((3, "mixed_ir"),))
        for precision, tag in hpcg_modes:
            xs = alp.Vector(n, fill=0.0)
            H.solve(b, xs, 2, 1e-10, precision)  # warm-up (graph capture, first touches)
            times = []
            for rep in range(3):
                xs.set(0.0)
                alp.sync()
                t0 = time.perf_counter()
                it, res, outer = H.solve(b, xs, 500, 1e-10, precision)
                times.append(time.perf_counter() - t0)
            r = alp.Vector(n)
            L.alp_hip_memcpy_d2d(r.ptr, b.ptr, b.nbytes)
            L.alp_hip_scale_f64(n, -1.0, r.ptr)
            alp.mxv(r, A, xs)
            true_rel = float(np.sqrt(alp.dot(r, r)) / bnorm)
            err = float(np.max(np.abs(xs.numpy() - 1.0)))
            ms = 1e3 * sorted(times)[1]
            solves[f"hpcg_{tag}"] = {"ms": round(ms, 2), "iterations": it, "refinements": outer,
                                     "rel_residual": true_rel, "max_err_vs_exact": err,
                                     "levels": H.levels, "colours": 8}
            log(f"[bench] HPCG {tag}: it={it} outer={outer} {ms:.1f} ms rel={true_rel:.2e} err={err:.2e}")
        for _, tag in hpcg_modes[1:]:
            solves[f"hpcg_speedup_{tag}_vs_fp64"] = round(solves["hpcg_fp64"]["ms"] / solves[f"hpcg_{tag}"]["ms"], 3)
        solves["hpcg_mixed_schemes"] = ("mixed = FP32 PCG iterations (FP32 V-cycle, FP64 row sums in A p, FP64 x) "
                                        "with reliable FP64 residual updates; mixed_pc = FP64 PCG with an FP32 "
                                        "V-cycle; mixed_ir = restarted FP32 PCG in FP64 iterative refinement")
        H.free()

    configs = {}
    if not args.no_configs:
        for name, fn in (("c1_cg50_64cubed", lambda: leg_c1(alp, L)), ("c2_rmat_s22", lambda: leg_c2(alp)),
                         ("c4_batched_64x8", lambda: leg_c4(alp)), ("c5_adjoint_128cubed", lambda: leg_c5(alp, L)),
                         ("grb_loops_64cubed", leg_grb_loops)):
            log(f"[bench] {name}")
            configs[name] = fn()
            log(f"[bench] {name}: {json.dumps(configs[name])}")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{side}^3 27-point stencil (n={n}, nnz={nnz}), FP64 plus-times "
                               f"grb::mxv y += A x on device-resident storage (CRS and the diagonal-slice copy the kernel streams)",
                   "kernel": KERNEL_NAMES.get(kernel, str(kernel)), "lanes": lanes, "rows": rows,
                   "bytes_per_step": fbytes},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "note": "achieved = the streamed layout's algorithmic bytes per launch "
                             "(format_bytes_per_launch) / the kernel's HIP-event time per launch; traffic = the "
                             "measured HBM bytes per launch (FETCH_SIZE + WRITE_SIZE, calibrated)"},
        "per_launch_us": per_launch,
        "format_bytes_per_launch": fbytes,
        "crs_bytes_per_launch": crs_bytes,
        "crs_equivalent_gbs": round(crs_equiv, 1),
        "crs_equivalent_note": "the metric's CRS bytes (12 nnz + 8 (n+1) + 8 n + 16 n) over the same kernel time; "
                               "not a fraction of peak, since the layout streams no column indices or row offsets",
        "crs_path": crs_path,
        "layout": {"slices": layout[0], "padded_entries": layout[1], "diagonal_slices": layout[2]},
        "build_ms": {"stencil_crs_sell_diagonal_f32": round(build_ms, 1),
                     "what": "generator, CRS, sliced ELL, diagonal slices and the FP32 copies, on the device"},
        "mxv_overwrite": {"us_per_launch": round(ow_ms * 1e3, 2), "gbs": round(ow_gbs, 1)},
        "solves": solves,
        "configs": configs,
    }
    tr = measured_traffic(kernel, lanes, rows)
    if tr is not None:
        out["roofline"]["traffic"] = tr[0]
        out["roofline"]["traffic_source"] = tr[1]
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline")
        out["cpu_baseline"] = cpu_baseline(side, nnz, n)
        if solves.get("cg_fp64") and solves.get("hpcg_fp64"):
            log("[bench] CPU reference solves (CG whole, HPCG sampled)")
            ref = cpu_solve_reference(side, solves["cg_fp64"]["iterations"], solves["hpcg_fp64"]["iterations"])
            if ref:
                if ref.get("cg_fp64"):
                    c = ref["cg_fp64"]["ms"]
                    ref["cg_fp64"]["device_fp64_speedup"] = round(c / solves["cg_fp64"]["ms"], 1)
                    if solves.get("cg_mixed"):
                        ref["cg_fp64"]["device_mixed_speedup"] = round(c / solves["cg_mixed"]["ms"], 1)
                out["cpu_baseline"]["solves"] = ref
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
