#!/usr/bin/env python3
"""Headline benchmark: mixed-precision MGPCG pressure solves of the 3D pure-Neumann 7-point Poisson system at 256^3
to 1e-8 true relative residual, on one MI355X.

A step is one complete solve (right-hand side already resident in HBM, solution left in HBM): scatter to the box,
FP64 residual, FP32 multigrid-preconditioned CG passes inside FP64 iterative refinement, mean projection, gather.
Prints one JSON line. Also reports, on the same GPU, the FP64 matrix-free Jacobi-PCG solve of the same system (the
plain-CG comparison), and on the host the reference's own AMGCL and Eigen CG (oracle/_ref, compiled from the
reference sources) on a bounded sample.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--size 256] [--no-cpu] [--no-cg64]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "MGPCG solves/s @256³ to 1e-8 rel res, HBM GB/s; p50/p99 solve latency"
SEED = 20261203


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(n):
    import nereus_cfd_amd as N
    g = N.uniform_grid(n)                       # all six faces SOLID: pure Neumann, singular
    b = N.synth_rhs(g, seed=SEED, modes=8, noise=0.01, zero_mean=True)
    return g, b


def gpu_leg(args):
    import torch
    import nereus_cfd_amd as N

    n = args.size
    g, b = workload(n)
    log(f"[bench] workload {n}^3 unknowns={len(b)}")
    dev = torch.device("cuda:0")
    tb = torch.from_numpy(b).to(dev)
    tx = torch.zeros_like(tb)
    torch.cuda.synchronize()

    s = N.Solver(method=N.METHOD_MGPCG, tolerance=1e-8, max_iters=2000, device=0)
    t0 = time.perf_counter()
    s.commit(g)
    setup_s = time.perf_counter() - t0
    log(f"[bench] commit {setup_s:.2f}s, device memory {s.memory_bytes() / 2**20:.0f} MiB")
    for w in range(args.warmup):
        res, rc = s.solve_device(tb.data_ptr(), tx.data_ptr())
        log(f"[bench] warmup {w}: iters={res.iterations} passes={res.refinements} relres={res.relative_residual:.3e} "
            f"device_ms={res.device_ms:.3f}")
    s.profile(True)
    lat = []
    bytes_total = 0.0
    iters = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        res, rc = s.solve_device(tb.data_ptr(), tx.data_ptr())
        lat.append(time.perf_counter() - ts)
        bytes_total += res.algorithmic_bytes
        iters.append(res.iterations)
        if rc != 0 or not res.relative_residual < 1e-8:
            raise SystemExit(f"solve {k} did not converge: rc={rc} relres={res.relative_residual}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = s.profile_read()
    s.profile(False)
    log(f"[bench] timed {args.steps} solves in {elapsed:.4f}s")

    # Accuracy of the last result: true FP64 residual (reported by the solve) and the pressure error against the
    # FP64 matrix-free CG solution of the same system (relative L2 after mean removal).
    x = tx.cpu().numpy()
    relres = float(res.relative_residual)
    out = {
        "value": args.steps / elapsed,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "gbps": bytes_total / elapsed / 1e9,
        "algorithmic_gb_per_solve": bytes_total / args.steps / 1e9,
        "iterations": int(np.median(iters)),
        "refinements": int(res.refinements),
        "levels": int(res.levels),
        "relative_residual": relres,
        "setup_s": setup_s,
        "kernels": {k: {"launches": v["launches"], "ms": round(v["ms"], 4),
                        "gbps": round(v["bytes"] / max(v["ms"], 1e-12) / 1e6, 1)} for k, v in prof.items()},
    }
    dom = max(prof.items(), key=lambda kv: kv[1]["ms"])
    achieved = dom[1]["bytes"] / dom[1]["ms"] / 1e6
    out["roofline"] = {"bound": "hbm", "kernel": dom[0], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                       "avg_launch_us": round(1e3 * dom[1]["ms"] / dom[1]["launches"], 2)}
    s.close()

    if not args.no_cg64:
        c = N.Solver(method=N.METHOD_CG64, tolerance=1e-8, max_iters=20000, device=0)
        c.commit(g)
        ty = torch.zeros_like(tb)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r2, rc2 = c.solve_device(tb.data_ptr(), ty.data_ptr())
        torch.cuda.synchronize()
        t_cg = time.perf_counter() - t0
        y = ty.cpu().numpy()
        err = np.linalg.norm((x - x.mean()) - (y - y.mean())) / np.linalg.norm(y - y.mean())
        out["fp64_cg"] = {"solves_per_s": 1.0 / t_cg, "seconds": t_cg, "iterations": int(r2.iterations),
                          "relative_residual": float(r2.relative_residual),
                          "gbps": r2.algorithmic_bytes / t_cg / 1e9,
                          "speedup_mgpcg_vs_fp64_cg": out["value"] * t_cg}
        out["pressure_error_vs_fp64_cg"] = float(err)
        c.close()
        log(f"[bench] fp64 CG: {t_cg:.3f}s iters={r2.iterations} relres={r2.relative_residual:.3e} err={err:.3e}")
    return out


def cpu_leg(args):
    """The reference's CPU solvers (compiled from its sources into oracle/_ref) on a bounded sample: AMGCL CG with
    smoothed aggregation (the reference's eigen_solver=AMG path) on the 128^3 member of the same workload family,
    one solve after hierarchy set-up, scaled to a 256^3 solve by the cell count (AMG work is linear in it)."""
    from oracle import refcase as R
    import nereus_cfd_amd as N
    if not R.available():
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    n = 128
    g, b = workload(n)
    case = R.RefCase(dim=3, grid_size=n, types=g.types, compat=True,
                     solves=[R.SolveSpec(R.AMG, R.DIAGONAL, 1e-8, 500, repeats=2),
                             R.SolveSpec(R.CG, R.DIAGONAL, 1e-8, 60, repeats=1)], rhs=b)
    t0 = time.perf_counter()
    out = R.run(case, threads=threads, timeout=600)
    wall = time.perf_counter() - t0
    amg = out["solves"][0]
    cg = out["solves"][1]
    t_amg = float(amg["times"][-1])           # second solve: hierarchy already built (AMG_CACHE)
    t_setup = float(amg["times"][0] - amg["times"][-1])
    scale = (256 / n) ** 3
    cg_per_iter = float(cg["times"][0]) / max(1, cg["iters"] + 1)
    return {
        "value": 1.0 / (t_amg * scale), "unit": "solves/s", "cores": out["threads"], "kind": "reference",
        "sample": f"reference AMGCL CG+SA/SPAI0 (PressureSettings.h:325-379) on the {n}^3 member of the workload to "
                  f"1e-8 ({amg['iters']} iterations, {t_amg:.3f}s after a {t_setup:.2f}s hierarchy set-up), "
                  f"scaled x{scale:.0f} to 256^3; whole CPU leg {wall:.1f}s",
        "eigen_cg_s_per_iter_128": cg_per_iter,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cg64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    res = gpu_leg(args)
    cpu = None if args.no_cpu else cpu_leg(args)
    line = {
        "metric": METRIC,
        "value": round(res.pop("value"), 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res.pop("ms_per_step"), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": f"3D pure-Neumann 7-point Poisson {args.size}^3 cells, h=1/{args.size}, all faces "
                               f"SOLID; RHS 8 cosine modes + 0.01 N(0,1), mt19937_64 seed {SEED}, zero mean; "
                               f"MGPCG FP32 V(2,2) RBGS + FP64 refinement to 1e-8 true relative residual",
                   "cells": args.size ** 3, "tolerance": 1e-8},
        "roofline": res.pop("roofline"),
        "cpu_baseline": cpu,
    }
    line.update(res)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
