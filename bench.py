#!/usr/bin/env python3
"""Benchmark of the MI355X numeric phase of symla's multifrontal LDL^T.

Default workload (one "step"): numeric factorization of the 40^3 7-point
Dirichlet Laplacian shifted by -0.5*I (n = 64,000, nested-dissection
ordering) followed by a 16-right-hand-side solve, with the matrix values and
the right-hand sides already resident in HBM.  Two variants are timed on the
same plan: FP32 fronts (FP32 MFMA Schur updates, FP64 pivots) followed by a
solve with FP64 iterative refinement, and the plain FP64 factorize + solve.
Symbolic analysis and plan creation (device pool, scatter maps) happen once,
outside the timed regions.  Both variants must reach backward error <= 1e-12
with the expected inertia.

Other workloads (`--config`): c1 (2D 5-point 100^2 - 0.3 I, one RHS), c3
(SQD KKT system in static-regularized mode, 4 RHS), c4 (a batch of
independent 2D systems with random couplings, one plan), c5 (27-point
stencil with random weights on 80^3 - 0.2 I, 64 RHS, plus replays of a
recorded FP32 refactorize + solve on new values).

Prints ONE JSON line.  `value` is seconds per FP32+IR factorize+solve (lower
is better); `fp64_s` is the plain FP64 step.  `roofline` covers the dominant
kernel of the FP32 variant (the FP32 MFMA trailing update) from HIP events
recorded by the library around each of its launches on the plan's stream
during the timed steps (`roofline_fp64`: the FP64 update of the FP64
variant).  Large fronts with big contribution blocks update those blocks
once, after their last panel, as library SYRKs: that part is reported
beside the roofline as `cb_update` (and `schur_total_tflops` covers both
kinds of update); `cpu_baseline` times the CPU oracle (a restatement of the
reference's OpenMP multifrontal path) on a bounded sample of the same
workload.
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

PEAK_FP64_MFMA_TFLOPS = 78.6   # MI355X dense FP64 matrix peak
PEAK_FP32_MFMA_TFLOPS = 157.3  # MI355X dense FP32 matrix peak
SEED = 0x53594D4C41
BE_BAR = 1e-12  # backward-error bar of the metric


def build_workload(cfg, args, P, OrderingType, Mode, SymLDLT, BatchedLDLT, Precision):
    """(description, A_lower, solver, expected inertia or None, nrhs, config dict)."""
    if cfg == "c1":
        A = P.laplace2d(100, 100, 0.3)
        s = SymLDLT()
        s.setOrdering(OrderingType.AMD)
        s.analyzePattern(A)
        exp = tuple(P.closed_form_inertia(P.laplace_eigs_2d(100, 100), 0.3)[:3])
        return ("2D 5-point Laplacian 100^2 - 0.3 I (n=10000), minimum-degree ordering", A, s, exp, 1, {})
    if cfg == "c2":
        G = args.grid
        A = P.laplace3d(G, G, G, 0.5)
        s = SymLDLT()
        s.setOrdering(OrderingType.Metis)
        s.analyzePattern(A)
        exp = tuple(P.closed_form_inertia(P.laplace_eigs_3d(G, G, G), 0.5)[:3])
        return (f"3D 7-point Laplacian {G}^3 - 0.5 I (n={A.shape[0]}), nested dissection", A, s, exp, args.nrhs, {})
    if cfg == "c3":
        K, n1, n2 = P.sqd_kkt(30, 9000)
        s = SymLDLT()
        s.setOrdering(OrderingType.AMD)
        s.setMode(Mode.StaticRegularized)
        s.setKKTBlockSizes(n1, n2)
        s.analyzePattern(K)
        return ("SQD KKT [[-(H+1e-4 I), J^T], [J, 1e-8 I]], H 7-point 30^3, J 9000 x 27000 (5 per row); "
                "static-regularized, minimum-degree ordering", K, s, (n2, n1, 0), 4, {"n1": n1, "n2": n2})
    if cfg == "c4":
        systems = [P.batched_system(SEED + 4 + 1000 * i) for i in range(args.systems)]
        b = BatchedLDLT(OrderingType.AMD)
        b.analyzePattern(systems)
        return (f"{args.systems} independent 2D 5-point systems + 2 random couplings per row, n_i ~ U[200, 5000], "
                f"- 0.3 I, one block-diagonal plan", b.matrix, b.solver, None, 1,
                {"systems": args.systems, "sum_n": int(b.matrix.shape[0])})
    if cfg == "c5":
        N = args.grid if args.grid != 40 else 80
        A = P.stencil27(N, 0.2)
        s = SymLDLT()
        s.setOrdering(OrderingType.Metis)
        s.analyzePattern(A)
        return (f"27-point stencil {N}^3, U(0.5, 1.5) edge weights, - 0.2 I (n={A.shape[0]}), nested dissection",
                A, s, None, args.nrhs if args.nrhs != 16 else 64, {"grid": N})
    raise SystemExit(f"unknown config {cfg}")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--grid", type=int, default=40)
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--systems", type=int, default=256, help="c4: number of independent systems")
    ap.add_argument("--replays", type=int, default=10, help="c5: recorded refactorize + solve replays")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: min(16, cpu count))")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy 256 MiB device to device (the known-size kernel that scales the rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--exact-panel", action="store_true",
                    help="panels without the speculative blocked step (pivot-by-pivot kernel only)")
    ap.add_argument("--spec-min-rows", type=int, default=None,
                    help="fronts with more eligible rows than this take the speculative panel step")
    ap.add_argument("--small-max", type=int, default=None,
                    help="largest front order factored whole in LDS (at most 192)")
    ap.add_argument("--cb-min-rows", type=int, default=None,
                    help="contribution blocks of at least this many rows are updated once after the front's last "
                         "panel (-1: panel by panel)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one MI355X: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import torch
    if args.calibrate_copy:
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": src.numel() * 8}))
        return 0

    from ma57_amd import _lib, problems as P
    from ma57_amd.batch import BatchedLDLT
    from ma57_amd.solver import SymLDLT, OrderingType, Mode, Precision, csc_arrays

    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        return 3
    dev = torch.device("cuda:0")
    lib = _lib.load()

    t0 = time.perf_counter()
    desc, A, solver, expect, nrhs, extra = build_workload(args.config, args, P, OrderingType, Mode, SymLDLT,
                                                          BatchedLDLT, Precision)
    analyze_s = time.perf_counter() - t0
    n = A.shape[0]
    sf = solver.symbolicFactor()
    plan = solver._plan
    if args.exact_panel or args.spec_min_rows is not None or args.cb_min_rows is not None or args.small_max is not None:
        small, tile = solver.tuning()
        if args.small_max is not None:
            small = args.small_max
        solver.setTuning(small, tile, not args.exact_panel, args.spec_min_rows, args.cb_min_rows)
    tuning = solver._get_tuning()
    B = np.random.default_rng(SEED + int(args.config[1])).standard_normal((n, nrhs))

    _, _, _, vals = csc_arrays(A)
    d_vals = torch.from_numpy(vals).to(dev)
    d_B = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)  # (nrhs, n) row-major == (n, nrhs) column-major
    d_X = torch.empty_like(d_B)
    torch.cuda.synchronize()
    front_flops = sum(s.ncols * len(s.rowPattern) ** 2 - s.ncols ** 2 * len(s.rowPattern) + s.ncols ** 3 / 3.0
                      for s in sf.supernodes)

    def leg(precision, refine):
        """Time `steps` factorize + solve steps of one precision; refine=True solves with FP64 refinement."""
        solver.setPrecision(precision)
        opts = solver._opts()
        signs = solver._expected_signs() if solver.mode == Mode.StaticRegularized else None
        if signs is not None:
            opts.expected_sign = signs.ctypes.data_as(_lib.i32p)
        info = _lib.FactorInfo()
        rinfo = _lib.RefineInfo()
        # refinement stops once the relative residual (the reference's stored-entry norm of A, which is no
        # larger than the symmetric one) reaches the 1e-12 bar; the reported backward error is recomputed
        # per column with the symmetric norm
        ropts = _lib.RefineOpts(10, BE_BAR)

        def step():
            rc = lib.symla_factorize_device(plan, C.c_void_p(d_vals.data_ptr()), C.byref(opts), C.byref(info))
            if rc:
                raise RuntimeError(lib.symla_plan_last_error(plan).decode())
            if refine:
                rc = lib.symla_solve_refined_device(plan, C.c_void_p(d_vals.data_ptr()), C.c_void_p(d_B.data_ptr()),
                                                    n, nrhs, C.c_void_p(d_X.data_ptr()), n, C.byref(ropts),
                                                    C.byref(rinfo))
            else:
                rc = lib.symla_solve_device(plan, C.c_void_p(d_B.data_ptr()), n, nrhs, C.c_void_p(d_X.data_ptr()), n)
            if rc:
                raise RuntimeError(lib.symla_plan_last_error(plan).decode())

        for _ in range(args.warmup):
            step()
        # timed region (kernel-class events recorded by the library on the plan's stream)
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
        # the same steps without event brackets
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        clean = (time.perf_counter() - t1) / args.steps
        fms, sms = C.c_double(), C.c_double()
        lib.symla_plan_timing(plan, C.byref(fms), C.byref(sms))
        X = d_X.cpu().numpy().T
        syrk_avg_ms = prof.syrk_ms / max(prof.syrk_launches, 1)
        flops_per_launch = prof.schur_flops / max(prof.syrk_launches, 1)
        achieved = flops_per_launch / (syrk_avg_ms * 1e-3) / 1e12 if syrk_avg_ms > 0 else 0.0
        return {
            "s": elapsed / args.steps,
            "s_without_event_brackets": clean,
            "backward_error": P.backward_error(A, X, B),
            "inertia": [info.n_pos, info.n_neg, info.n_zero],
            "ir_iterations": rinfo.iterations if refine else 0,
            "static_retries": info.retries_used,
            "device_factor_ms": fms.value,
            "kernel_ms_per_step": {k: getattr(prof, k + "_ms") / args.steps for k in
                                   ("syrk", "cb", "panel", "panel_rows", "small", "assemble", "fwd", "bwd")},
            "syrk": {"achieved": achieved, "launches_per_step": prof.syrk_launches / args.steps,
                     "avg_launch_ms": syrk_avg_ms, "flops_per_launch": flops_per_launch,
                     "schur_gflop_per_step": prof.schur_flops / args.steps / 1e9},
            # deferred contribution-block updates: operand kernel + library SYRK per chunk
            "cb_update": {"achieved": (prof.cb_flops / (prof.cb_ms * 1e-3) / 1e12) if prof.cb_ms > 0 else 0.0,
                          "launches_per_step": prof.cb_launches / args.steps,
                          "gflop_per_step": prof.cb_flops / args.steps / 1e9},
            # every Schur-complement flop over the time of both kinds of update
            "schur_total_tflops": ((prof.schur_flops + prof.cb_flops) / ((prof.syrk_ms + prof.cb_ms) * 1e-3) / 1e12
                                   if prof.syrk_ms + prof.cb_ms > 0 else 0.0),
        }

    # the plain FP64 variant solves without refinement, except where the matrix needs it to reach the
    # 1e-12 bar (static perturbations on c3; the 27-point system's conditioning on c5)
    fp64 = leg(Precision.FP64, refine=args.config in ("c3", "c5"))
    mixed = leg(Precision.FP32, refine=True)
    if expect is None:
        expect = tuple(fp64["inertia"])  # no closed form: the FP64 factorization of the same matrix

    def traffic(kernel):
        """HBM bytes per dispatch from the committed counter passes of the same workload, or None."""
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic_schur_laplace3d_40.json")
        if args.config != "c2" or nrhs != 16 or not os.path.exists(path):
            return None
        with open(path) as fh:
            rec = json.load(fh)
        k = rec.get("kernels", {}).get(kernel)
        return k["hbm_bytes_per_dispatch"] if k else None

    def roofline(kernel, r, peak):
        return {"bound": "mfma", "kernel": kernel, "achieved": r["syrk"]["achieved"], "peak": peak,
                "unit": "TFLOP/s", "frac": r["syrk"]["achieved"] / peak, "traffic": traffic(kernel),
                "launches_per_step": r["syrk"]["launches_per_step"], "avg_launch_ms": r["syrk"]["avg_launch_ms"],
                "flops_per_launch": r["syrk"]["flops_per_launch"]}

    graph = None
    if args.config == "c5" and args.replays > 0:
        # recorded FP32 refactorize + solve, replayed on new seeded values of the same pattern
        N = extra["grid"]
        value_sets = [torch.from_numpy(csc_arrays(P.stencil27(N, 0.2, seed=SEED + 5 + r))[3]).to(dev)
                      for r in range(args.replays)]
        solver.setPrecision(Precision.FP32)
        solver.captureGraph(d_vals.data_ptr(), d_B.data_ptr(), nrhs, d_X.data_ptr())
        times, bes = [], []
        for r, vs in enumerate(value_sets):
            d_vals.copy_(vs)
            torch.cuda.synchronize()
            tg = time.perf_counter()
            solver.replayGraph()
            torch.cuda.synchronize()
            times.append(time.perf_counter() - tg)
            if r == 0:
                Ar = P.stencil27(N, 0.2, seed=SEED + 5)
                bes.append(P.backward_error(Ar, d_X.cpu().numpy().T, B))
        solver.releaseGraph()
        graph = {"replays": args.replays, "s_per_replay": float(np.mean(times)),
                 "s_per_replay_min": float(np.min(times)), "fp32_unrefined_backward_error_first": bes[0]}

    cpu = None
    if not args.no_cpu_baseline and args.config in ("c1", "c2"):
        from oracle import oracle as O

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        sym = O.Symbolic.from_product(sf)
        tc = time.perf_counter()
        f = O.mf_factorize(sym, A, parallel=True, threads=threads)
        Xc = f.solve(B, parallel=True, threads=threads)
        cpu_s = time.perf_counter() - tc
        cpu = {"value": cpu_s, "unit": "s", "cores": threads, "kind": "port",
               "sample": f"one full FP64 factorize + {nrhs}-RHS solve of the same system (OpenMP task DAG "
                         f"restatement of the reference path, {threads} threads)",
               "inertia": list(f.inertia()), "backward_error": P.backward_error(A, Xc, B)}

    ok64 = tuple(fp64["inertia"]) == tuple(expect) and fp64["backward_error"] <= 1e-12
    ok32 = tuple(mixed["inertia"]) == tuple(expect) and mixed["backward_error"] <= 1e-12
    out = {
        "metric": "factorize+solve s, FP32+IR vs FP64 at backward err <=1e-12; Schur TFLOP/s",
        "value": mixed["s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["s"] * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": dict({"workload": f"{desc}; factorize + {nrhs}-RHS solve; value: FP32 fronts + FP64 iterative "
                                    f"refinement, fp64: plain FP64 factorize + solve",
                        "name": args.config, "n": n, "nrhs": nrhs, "nnz_L": int(sf.nnz_l),
                        "supernodes": len(sf.supernodes),
                        "max_front": max(len(s.rowPattern) for s in sf.supernodes),
                        "small_front_max": tuning.small_front_max,
                        "speculative_panel": bool(tuning.speculative_panel),
                        "speculative_min_rows": tuning.speculative_min_rows,
                        "deferred_cb_min_rows": tuning.deferred_cb_min_rows}, **extra),
        "roofline": roofline("k_schur<float>", mixed, PEAK_FP32_MFMA_TFLOPS),
        "cpu_baseline": cpu,
        "fp32_ir": mixed,
        "fp64": fp64,
        "fp64_s": fp64["s"],
        "speedup_fp32ir_vs_fp64": fp64["s"] / mixed["s"],
        "roofline_fp64": roofline("k_schur<double>", fp64, PEAK_FP64_MFMA_TFLOPS),
        "backward_error": mixed["backward_error"],
        "inertia": mixed["inertia"],
        "inertia_expected": list(expect),
        "ir_iterations": mixed["ir_iterations"],
        "analyze_s": analyze_s,
        "factor_gflop": front_flops / 1e9,
        "graph": graph,
    }
    print(json.dumps(out))
    return 0 if (ok64 and ok32) else 1


if __name__ == "__main__":
    sys.exit(main())
