#!/usr/bin/env python3
"""Benchmark of the MI355X numeric phase on the 3D 7-point indefinite Laplacian.

Workload (one "step"): numeric factorization of the 40^3 7-point Dirichlet
Laplacian shifted by -0.5*I (n = 64,000, nested-dissection ordering) followed
by a 16-right-hand-side solve, FP64 throughout, with the matrix values and the
right-hand sides already resident in HBM.  Symbolic analysis and plan
creation (device pool, scatter maps) happen once, outside the timed region.

Prints ONE JSON line.  `value` is seconds per factorize+solve (lower is
better).  `roofline` covers the dominant kernel (the FP64 MFMA trailing
update) from HIP events recorded by the library around each of its launches
on the plan's stream during the timed steps; `cpu_baseline` times the CPU
oracle (a restatement of the reference's OpenMP multifrontal path) on the
same workload.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA_TFLOPS = 78.6  # MI355X dense FP64 matrix peak
PEAK_HBM_GBS = 8000.0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--grid", type=int, default=40)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: min(16, cpu count))")
    ap.add_argument("--verbose", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import scipy.sparse as sp  # noqa: F401
    import torch

    from ma57_amd import _lib, problems as P
    from ma57_amd.solver import SymLDLT, OrderingType, csc_arrays

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 3
    dev = torch.device("cuda:0")
    lib = _lib.load()

    G = args.grid
    shift = 0.5
    A = P.laplace3d(G, G, G, shift)
    n = A.shape[0]
    rng = np.random.default_rng(0x53594D4C41 + 2)
    B = rng.standard_normal((n, args.nrhs))

    t0 = time.perf_counter()
    solver = SymLDLT()
    solver.setOrdering(OrderingType.Metis)
    solver.analyzePattern(A)
    analyze_s = time.perf_counter() - t0
    sf = solver.symbolicFactor()
    plan = solver._plan

    _, _, _, vals = csc_arrays(A)
    d_vals = torch.from_numpy(vals).to(dev)
    d_B = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)  # (nrhs, n) row-major == (n, nrhs) column-major
    d_X = torch.empty_like(d_B)
    torch.cuda.synchronize()
    opts = solver._opts()
    info = _lib.FactorInfo()

    def step():
        rc = lib.symla_factorize_device(plan, C.c_void_p(d_vals.data_ptr()), C.byref(opts), C.byref(info))
        if rc:
            raise RuntimeError(lib.symla_plan_last_error(plan).decode())
        rc = lib.symla_solve_device(plan, C.c_void_p(d_B.data_ptr()), n, args.nrhs, C.c_void_p(d_X.data_ptr()), n)
        if rc:
            raise RuntimeError(lib.symla_plan_last_error(plan).decode())

    for _ in range(args.warmup):
        step()

    # timed region (kernel-class events recorded on the plan's stream)
    lib.symla_plan_set_profiling(plan, 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = _lib.Profile()
    lib.symla_plan_get_profile(plan, C.byref(prof))
    lib.symla_plan_set_profiling(plan, 0)

    # same steps without event brackets
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    clean = (time.perf_counter() - t1) / args.steps

    fms = C.c_double()
    sms = C.c_double()
    lib.symla_plan_timing(plan, C.byref(fms), C.byref(sms))

    X = d_X.cpu().numpy().T
    be = P.backward_error(A, X, B)
    inertia = (info.n_pos, info.n_neg, info.n_zero)
    expect = P.closed_form_inertia(P.laplace_eigs_3d(G, G, G), shift)[:3]

    ms_per_step = elapsed / args.steps * 1e3
    syrk_avg_ms = prof.syrk_ms / max(prof.syrk_launches, 1)
    flops_per_launch = prof.schur_flops / max(prof.syrk_launches, 1)
    achieved = flops_per_launch / (syrk_avg_ms * 1e-3) / 1e12 if syrk_avg_ms > 0 else 0.0

    front_flops = sum(s.ncols * len(s.rowPattern) ** 2 - s.ncols ** 2 * len(s.rowPattern) + s.ncols ** 3 / 3.0
                      for s in sf.supernodes)

    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        sym = O.Symbolic.from_product(sf)
        tc = time.perf_counter()
        f = O.mf_factorize(sym, A, parallel=True, threads=threads)
        Xc = f.solve(B, parallel=True, threads=threads)
        cpu_s = time.perf_counter() - tc
        cpu = {"value": cpu_s, "unit": "s", "cores": threads, "kind": "port",
               "sample": f"one full factorize + {args.nrhs}-RHS solve of the same {G}^3 system (OpenMP task DAG "
                         f"restatement of the reference path, {threads} threads)",
               "inertia": list(f.inertia()), "backward_error": P.backward_error(A, Xc, B)}

    out = {
        "metric": "factorize+solve s, FP32+IR vs FP64 at backward err <=1e-12; Schur TFLOP/s",
        "value": ms_per_step / 1e3,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "ms_per_step_without_event_brackets": clean * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D 7-point Laplacian {G}^3 - {shift} I (n={n}), nested dissection, "
                               f"FP64 factorize + solve with nrhs={args.nrhs}",
                   "n": n, "nrhs": args.nrhs, "precision": "fp64", "nnz_L": int(sf.nnz_l),
                   "supernodes": len(sf.supernodes), "max_front": max(len(s.rowPattern) for s in sf.supernodes)},
        "roofline": {"bound": "mfma", "kernel": "k_syrk_f64", "achieved": achieved,
                     "peak": PEAK_FP64_MFMA_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_MFMA_TFLOPS,
                     "traffic": None, "launches_per_step": prof.syrk_launches / args.steps,
                     "avg_launch_ms": syrk_avg_ms, "flops_per_launch": flops_per_launch},
        "cpu_baseline": cpu,
        "backward_error": be,
        "inertia": list(inertia),
        "inertia_expected": list(expect),
        "ir_iterations": 0,
        "analyze_s": analyze_s,
        "device_factor_ms": fms.value,
        "device_solve_ms": sms.value,
        "factor_gflop": front_flops / 1e9,
        "kernel_ms_per_step": {k: getattr(prof, k + "_ms") / args.steps for k in
                               ("syrk", "panel", "panel_rows", "small", "assemble", "fwd", "bwd")},
    }
    print(json.dumps(out))
    return 0 if (be <= 1e-12 and tuple(inertia) == tuple(expect)) else 1


if __name__ == "__main__":
    sys.exit(main())
