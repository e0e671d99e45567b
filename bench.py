#!/usr/bin/env python3
"""Benchmark of the MI355X SIMPLE hot path (one GPU).

Workloads
  psolve  pressure-equation solve to ||r||/||b|| <= 1e-10 on a 128^3 hex box
          with interior vertices jittered by +-20% h (seed 20261203),
          homogeneous Neumann everywhere except fixed_value 0 on x = 0,
          b ~ N(0,1) projected to zero mean, x0 = 0; FP64 PCG+AMG and
          FP32-inner / FP64-refined.  One step = one solve.

The dominant kernel's roofline is measured live: every Krylov SpMV of the
timed solves is bracketed by HIP events on the solver's stream; algorithmic
bytes per SpMV = 12 nnz + 4 (nC+1) + 16 nC (FP64 values + int32 columns,
row offsets, x read once, y written once).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 20261203


def psolve_problem(n):
    from anabasis_amd import libpoisson as lp
    m = lp.Mesh.hexbox(n, n, n, 1.0, 1.0, 1.0, jitter=0.2, seed=SEED)
    pat = lp.build_scalar_pattern(m)
    nb = m.nBoundaryFaces
    t = (m.bPatch[m.nInternalFaces:] == 1).astype(np.int32)  # xmin fixed_value 0
    v = np.zeros(nb)
    vals, _ = lp.assemble_scalar_elliptic_system(m, np.ones(m.nFaces), np.zeros(m.nCells), t, v,
                                                 np.zeros((m.nCells, 3)), False, nnz=pat.nnz)
    rng = np.random.default_rng(SEED)
    b = rng.standard_normal(m.nCells)
    b -= b.mean()
    return m, pat, vals, b


def run_psolve(args):
    from anabasis_amd import _lib as L
    from anabasis_amd import libpoisson as lp
    L.check(L.lib().an_set_device(0))
    n = args.n
    t0 = time.time()
    m, pat, vals, b = psolve_problem(n)
    print(f"# mesh {n}^3: nC={m.nCells} nnz={pat.nnz} build {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    res = {}
    for prec, name in ((0, "fp64"), (1, "mixed")):
        S = lp.PCGSolver(pat.rowOffsets, pat.cols, relTol=1e-10, absTol=0.0, precision=prec,
                         profile=1)
        ts = time.time()
        S.set_values(vals, rebuild=True)
        setup = time.time() - ts
        for _ in range(args.warmup):
            S.solve(b)
        times, infos = [], []
        for _ in range(args.steps):
            t1 = time.time()
            x, info = S.solve(b)
            times.append(time.time() - t1)
            infos.append(info)
        r = b - sp_mv(pat, vals, x)
        res[name] = dict(setup_s=setup, solve_s=float(np.mean(times)),
                         iterations=infos[-1].iterations,
                         rel_res=float(np.linalg.norm(r) / np.linalg.norm(b)),
                         refinement_steps=infos[-1].refinementSteps,
                         spmv_avg_s=float(np.mean([i.spmvAvgSeconds for i in infos])),
                         vcycle_avg_s=float(np.mean([i.vcycleAvgSeconds for i in infos])),
                         x=x)
        print(f"# {name}: setup {setup:.3f}s solve {res[name]['solve_s']:.4f}s "
              f"its {res[name]['iterations']} relres {res[name]['rel_res']:.2e} "
              f"spmv {res[name]['spmv_avg_s'] * 1e6:.1f}us vcycle "
              f"{res[name]['vcycle_avg_s'] * 1e6:.1f}us", file=sys.stderr, flush=True)
        del S
    nC, nnz = m.nCells, pat.nnz
    alg = 12 * nnz + 4 * (nC + 1) + 16 * nC
    t_sp = res["fp64"]["spmv_avg_s"]
    achieved = alg / t_sp / 1e9
    err_mixed = float(np.linalg.norm(res["mixed"]["x"] - res["fp64"]["x"]) /
                      np.linalg.norm(res["fp64"]["x"]))
    cpu = None if args.no_cpu_baseline else cpu_baseline_psolve(pat, vals, b, args)
    out = {
        "metric": "pressure solve time to 1e-10 relative residual (FP64 PCG+AMG), 128^3 "
                  "jittered hex",
        "value": res["fp64"]["solve_s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["fp64"]["solve_s"] * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"pressure solve, {n}^3 hex box jitter 0.2, Neumann + "
                               f"fixed_value x=0, b~N(0,1) mean-projected, rel tol 1e-10",
                   "cells": nC, "nnz": nnz},
        "psolve": {k: {kk: vv for kk, vv in v.items() if kk != "x"} for k, v in res.items()},
        "mixed_vs_fp64_rel_diff": err_mixed,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_spmv<double,double,double,0,1> (Krylov SpMV + fused p.q)",
                     "alg_bytes_per_launch": alg},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def sp_mv(pat, vals, x):
    import scipy.sparse as sp
    return sp.csr_matrix((vals, pat.cols, pat.rowOffsets)) @ x


def cpu_baseline_psolve(pat, vals, b, args):
    from oracle import pyoracle as O
    t0 = time.time()
    x, it, rr, _ = O.pcg_amg(pat.rowOffsets, pat.cols, vals, b, relTol=1e-10, absTol=0.0)
    dt = time.time() - t0
    print(f"# cpu oracle: {dt:.1f}s its {it}", file=sys.stderr, flush=True)
    return {"value": dt, "unit": "s", "cores": 1, "kind": "port",
            "sample": f"full solve (setup+solve) of the same {pat.nRows}-cell system by the "
                      "serial C oracle (same PCG + aggregation AMG)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="psolve", choices=["psolve"])
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X "
              "(multi-GPU sharding is described in DESIGN.md, not implemented)", file=sys.stderr)
        sys.exit(2)
    run_psolve(args)


if __name__ == "__main__":
    main()
