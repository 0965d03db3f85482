#!/usr/bin/env python3
"""Benchmark of the hip backend's hot path on one MI355X.

Workload: the 192^3 27-point stencil (7,077,888 rows, 189,119,224 nonzeros;
diagonal 26, off-diagonals -1), stored as device CRS. One step is one
grb::mxv<dense>(y, A, x) under the FP64 plus-times semiring, i.e. y = y + A x
(ALP's accumulate semantics, reference/blas2.hpp:1946).

value = SpMV throughput in GB/s of algorithmic bytes per step:
    12 nnz (8 B value + 4 B column index) + 8 (n+1) (row offsets)
    + 8 n (x, read once) + 16 n (y read + written)  = 2,495,923,112 B.
The solve legs (FP64 and mixed-precision CG / HPCG to 1e-10 relative residual)
are reported beside it. Inputs are resident in HBM before the timed region.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
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


def spmv_bytes(n: int, nnz: int, accumulate: bool = True) -> int:
    return 12 * nnz + 8 * (n + 1) + 8 * n + (16 if accumulate else 8) * n


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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--side", type=int, default=192)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--kernel", type=int, default=0,
                    help="skip the autotuner: 1 = CSR-vector, 2 = CSR-stream (with --lanes/--rows)")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a plain streaming kernel of known size (f64 -> f32 conversion of "
                         "2^28 elements: 2 GiB read, 1 GiB written) for HBM counter calibration")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md); run with --gpus 1", file=sys.stderr)
        return 2

    import alp_amd as alp
    alp.init(0)
    L = alp.lib()
    side = args.side
    log(f"[bench] building {side}^3 27-point stencil on the device")
    A = alp.Matrix.stencil27(side, side, side, flags=alp.MAT_WITH_F32)
    n, nnz = A.nrows, A.nnz
    if args.kernel:
        A.set_kernel(args.kernel, args.lanes, args.rows)
        kernel, lanes, rows = A.get_kernel()
    else:
        kernel, lanes, rows = A.autotune()
    log(f"[bench] n={n} nnz={nnz} kernel={'csr-vector' if kernel == 1 else 'csr-stream'} lanes={lanes} rows={rows}")
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
    bytes_step = spmv_bytes(n, nnz)
    achieved = bytes_step / (kern_ms * 1e-3) / 1e9
    value = bytes_step / (ms_step * 1e-3) / 1e9
    log(f"[bench] mxv: {kern_ms * 1e3:.1f} us/launch (events), {ms_step * 1e3:.1f} us/step (wall), "
        f"{achieved:.0f} GB/s")

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
        for precision, tag in ((0, "fp64"), (1, "mixed")):
            xs = alp.Vector(n, fill=0.0)
            alp.sync()
            t0 = time.perf_counter()
            rc, it, res = alp.cg(A, b, xs, 20000, 1e-10, precision)
            dt = time.perf_counter() - t0
            err = float(np.max(np.abs(xs.numpy() - 1.0)))
            solves[f"cg_{tag}"] = {"ms": round(dt * 1e3, 2), "iterations": it, "rc": rc,
                                   "rel_residual": float(np.sqrt(res) / np.sqrt(float(alp.dot(b, b)))),
                                   "max_err_vs_exact": err}
            log(f"[bench] CG {tag}: rc={rc} it={it} {dt * 1e3:.1f} ms err={err:.2e}")
        if "cg_fp64" in solves and "cg_mixed" in solves:
            solves["cg_speedup_mixed_vs_fp64"] = round(solves["cg_fp64"]["ms"] / solves["cg_mixed"]["ms"], 3)
        # HPCG: 4 grids (192/96/48/24), 8-colour symmetric Gauss-Seidel, b = A*1, x0 = 0
        log("[bench] building the HPCG hierarchy")
        H = alp.Hpcg(side, side, side, max_levels=3, colours=0)
        bnorm = np.sqrt(float(alp.dot(b, b)))
        for precision, tag in ((0, "fp64"), (1, "mixed_ir"), (2, "mixed_pc")):
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
        for tag in ("mixed_ir", "mixed_pc"):
            solves[f"hpcg_speedup_{tag}_vs_fp64"] = round(solves["hpcg_fp64"]["ms"] / solves[f"hpcg_{tag}"]["ms"], 3)
        H.free()

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
                               f"grb::mxv y += A x on device CRS",
                   "kernel": "csr-vector" if kernel == 1 else "csr-stream", "lanes": lanes, "rows": rows,
                   "bytes_per_step": bytes_step},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "mxv_overwrite": {"us_per_launch": round(ow_ms * 1e3, 2), "gbs": round(ow_gbs, 1)},
        "solves": solves,
    }
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline")
        out["cpu_baseline"] = cpu_baseline(side, nnz, n)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
