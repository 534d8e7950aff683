#!/usr/bin/env python3
"""Headline benchmark: mixed-precision MGPCG pressure solves of the 3D pure-Neumann 7-point Poisson system at 256^3
to 1e-8 true relative residual, on one MI355X.

A step is one complete solve (right-hand side already resident in HBM, solution left in HBM): scatter to the box
(with no guess the residual is b itself), CG with its recursion in FP64 preconditioned by an FP32 multigrid
V-cycle, the true FP64 residual, mean projection, gather. The dominant kernel's launch durations come from HIP
events that its own dispatches record (start and end of each finest-level sweep) in every tenth timed solve.
Prints one JSON line. Also reports, on the same GPU, the FP64 matrix-free Jacobi-PCG solve of the same system (the
plain-CG comparison), and on the host the reference's own AMGCL and Eigen CG (oracle/_ref, compiled from the
reference sources) on a bounded sample.

    python bench.py [--gpus 1] [--steps 50] [--warmup 2] [--size 256] [--no-cpu] [--no-cg64]
    python bench.py --config channel|batched|serving|step|layout   (the other configurations, one JSON line each)
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


def measured_traffic():
    """HBM bytes per launch of the finest-level sweep at 256^3, from the FETCH_SIZE and WRITE_SIZE counter runs
    summarised (with their commands and the copy-kernel calibration) in profiles/k_rbgs_row_256cube_hbm_traffic.md;
    None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_rbgs_row_256cube_hbm_traffic.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch ="):
                    return int(line.split("=")[1])
    except OSError:
        pass
    return None


def profiled_launch_us():
    """The dominant kernel's average duration per launch in the 256^3 kernel trace summarised in
    profiles/k_rbgs_row_256cube_mgpcg_kernel_trace.md (rocprofv3 --kernel-trace of this benchmark), as a cross-check
    of the live figure; None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_rbgs_row_256cube_mgpcg_kernel_trace.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("rocprof_avg_launch_us ="):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


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
    # In every tenth timed solve (the last of each ten) each finest-level red-black sweep (the dominant kernel)
    # records a pair of HIP events with its own dispatch (hipExtLaunchKernel's start and stop events, on the
    # solver's stream): the kernel's duration, as a kernel trace measures it. The other solves run without events
    # (a profiled solve launches its iterations directly instead of replaying the captured graph).
    every = 10 if args.steps >= 10 else args.steps
    dom_launches, dom_ms, dom_bytes = 0, 0.0, 0.0
    lat = []
    bytes_total = 0.0
    iters = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        sampled = k % every == every - 1
        if sampled:
            s.profile(True, kinds=["smooth0"])
        ts = time.perf_counter()
        res, rc = s.solve_device(tb.data_ptr(), tx.data_ptr())
        lat.append(time.perf_counter() - ts)
        if sampled:
            p0 = s.profile_read().get("smooth0", {"launches": 0, "ms": 0.0, "bytes": 0.0})
            s.profile(False)
            dom_launches += p0["launches"]
            dom_ms += p0["ms"]
            dom_bytes += p0["bytes"]
        bytes_total += res.algorithmic_bytes
        iters.append(res.iterations)
        if rc != 0 or not res.relative_residual < 1e-8:
            raise SystemExit(f"solve {k} did not converge: rc={rc} relres={res.relative_residual}")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    dom = ("smooth0", {"launches": dom_launches, "ms": dom_ms, "bytes": dom_bytes})
    log(f"[bench] timed {args.steps} solves in {elapsed:.4f}s")
    # Per-kind breakdown of one further (untimed) solve with events around every launch.
    s.profile(True)
    s.solve_device(tb.data_ptr(), tx.data_ptr())
    prof = s.profile_read()
    s.profile(False)

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
    achieved = dom[1]["bytes"] / dom[1]["ms"] / 1e6
    traffic = measured_traffic() if n == 256 else None
    out["roofline"] = {"bound": "hbm", "kernel": "k_rbgs_row (finest-level red-black Gauss-Seidel sweep)",
                       "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                       "launches": dom[1]["launches"],
                       "avg_launch_us": round(1e3 * dom[1]["ms"] / dom[1]["launches"], 2),
                       "bytes_per_launch": round(dom[1]["bytes"] / dom[1]["launches"]),
                       "timing": "HIP start/stop events recorded by each sampled launch's own dispatch "
                                 "(hipExtLaunchKernel) on the solver's stream",
                       "rocprof_avg_launch_us": profiled_launch_us() if n == 256 else None,
                       # bytes_per_launch follows the sweep formula (3s+f)N, one code byte per box cell included;
                       # every row of this box has a code descriptor, so the sweeps do not read those N bytes
                       "bytes_moved_per_launch": round(dom[1]["bytes"] / dom[1]["launches"] - (n + 2) ** 3),
                       "frac_bytes_moved": round((dom[1]["bytes"] / dom[1]["launches"] - (n + 2) ** 3)
                                                 / (dom[1]["bytes"] / dom[1]["launches"]) * achieved
                                                 / HBM_PEAK_GBS, 4)}
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


def host_info():
    """The host the CPU baseline runs on: CPU model, logical CPUs, the CPUs this process may use, physical cores."""
    model, cores = "unknown", set()
    try:
        phys = core = None
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None and core is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_allowed": allowed,
            "physical_cores": len(cores) or None}


def cpu_leg(args):
    """The reference's CPU solvers (compiled from its sources into oracle/_ref) on the headline system itself
    (PressureSettings.h:325-379, :400): AMGCL CG with smoothed aggregation and SPAI0 (eigen_solver=AMG) to 1e-8,
    its hierarchy set-up timed apart (the first solve builds it, AMG_CACHE keeps it), then five solves from x0 = 0
    with the hierarchy reused: value = 1 / their median; and Eigen's Jacobi CG (eigen_solver=CG) run for a bounded
    40 iterations, reported per iteration and projected to the iteration count the same CG needs (the FP64 CG on the
    GPU follows Eigen's iteration semantics). Threads: OMP_NUM_THREADS, else the CPUs this process may use."""
    from oracle import refcase as R
    if not R.available():
        return None
    info = host_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or info["cpus_allowed"]
    n = args.size
    g, b = workload(n)
    reps = 6
    case = R.RefCase(dim=3, grid_size=n, types=g.types, compat=True,
                     solves=[R.SolveSpec(R.AMG, R.DIAGONAL, 1e-8, 500, repeats=reps),
                             R.SolveSpec(R.CG, R.DIAGONAL, 1e-8, 40, repeats=1)], rhs=b,
                     write_matrix=False, write_solutions=False)
    t0 = time.perf_counter()
    out = R.run(case, threads=threads, timeout=900)
    wall = time.perf_counter() - t0
    amg, cg = out["solves"][0], out["solves"][1]
    times = np.asarray(amg["times"], float)
    solves = times[1:]
    med = float(np.median(solves))
    setup = float(times[0] - med)
    cg_per_iter = float(cg["times"][0]) / max(1, cg["iters"] + 1)
    log(f"[bench] cpu: AMGCL {n}^3 setup {setup:.2f}s, solves {np.round(solves, 3).tolist()} s "
        f"({amg['iters']} iterations), Eigen CG {cg_per_iter * 1e3:.1f} ms/iteration, {threads} threads, leg {wall:.1f}s")
    return {
        "value": 1.0 / med, "unit": "solves/s", "cores": out["threads"], "kind": "reference",
        "sample": f"reference AMGCL CG+SA/SPAI0 (PressureSettings.h:325-379) on the same {n}^3 system to 1e-8: "
                  f"{amg['iters']} iterations, median of {reps - 1} solves {med:.3f}s after a {setup:.2f}s hierarchy "
                  f"set-up (timed apart); whole CPU leg {wall:.1f}s",
        "p50_s": med, "p99_s": float(np.percentile(solves, 99)), "setup_s": setup, "amg_iterations": int(amg["iters"]),
        "eigen_cg_s_per_iter": cg_per_iter, "eigen_cg_sample_iterations": int(cg["iters"]) + 1,
        "threads": int(out["threads"]), **info,
    }


def channel_leg(args):
    """Stretched 256 x 128 x 128 channel with a voxelised cylinder and a p = 0 outlet, to 1e-10."""
    import torch
    import nereus_cfd_amd as N
    import nereus_cfd_amd.workloads as W
    g = W.stretched_channel()
    rng = np.random.default_rng(SEED)
    b = W.random_rhs(g, rng, zero_mean=False)
    tb = torch.from_numpy(b).cuda()
    tx = torch.zeros_like(tb)
    torch.cuda.synchronize()
    with N.Solver(tolerance=1e-10, max_iters=4000) as s:
        s.commit(g)
        for _ in range(args.warmup):
            s.solve_device(tb.data_ptr(), tx.data_ptr())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        lat = []
        for _ in range(args.steps):
            ts = time.perf_counter()
            res, rc = s.solve_device(tb.data_ptr(), tx.data_ptr())
            lat.append(time.perf_counter() - ts)
            if rc != 0 or not res.relative_residual < 1e-10:
                raise SystemExit(f"channel solve did not converge: {res.relative_residual}")
        el = time.perf_counter() - t0
    return {"metric": "stretched-channel MGPCG solves/s to 1e-10 rel res", "value": args.steps / el,
            "unit": "solves/s", "ms_per_step": 1e3 * el / args.steps, "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)), "iterations": int(res.iterations),
            "refinements": int(res.refinements), "relative_residual": float(res.relative_residual),
            "gbps": res.algorithmic_bytes / (el / args.steps) / 1e9,
            "config": {"workload": "3D stretched channel 256x128x128 cells, neighbour ratio <= 1.05 towards a "
                                   "cylinder r=0.05 (SOLID voxels), inflow x-min, p=0 outlet x-max, SOLID walls; "
                                   f"N(0,1) RHS seed {SEED}; MGPCG (V(4,4) FP32 preconditioner, FP64 CG "
                                   f"recursion) to 1e-10",
                       "unknowns": int(len(b))}}


def copy_leg(args):
    """A plain device copy of 1 GiB of FP32 (y = 1.0 * x, one read and one write of every byte, 4x the 256 MB
    Infinity Cache): the measured copy bandwidth of this GPU, and the known-size kernel that scales the HBM
    counters (FETCH_SIZE, WRITE_SIZE) of a profiled run to bytes."""
    import torch
    n = 1 << 28
    x = torch.rand(n, device="cuda:0", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(max(1, args.warmup)):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(args.steps):
        torch.mul(x, 1.0, out=y)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / args.steps
    moved = 2 * 4 * n
    return {"metric": "device copy GB/s", "value": round(moved / ms / 1e6, 1), "unit": "GB/s", "ms_per_step": ms,
            "scaling": "weak", "config": {"workload": "y = 1.0 * x over 2^28 FP32 (1 GiB read + 1 GiB write)",
                                          "bytes_read": 4 * n, "bytes_written": 4 * n}}


def step_leg(args):
    """The pressure step around the solve at 256^3 on MAC fields resident in HBM: a box with SOLID walls and a p = 0
    (EMPTY) outlet on the x-max ring, N(0,1) velocity; the right-hand side kernel, the projection kernel, and the
    composed pressure step (right-hand side, solve to 1e-8, scatter with the ring BC) plus projection step. Also the
    host-array right-hand side and projection (PCIe-inclusive) and the oracle's C loops on the host."""
    import torch
    import nereus_cfd_amd as N
    n = args.size
    g = N.uniform_grid(n)
    g.types[1:-1, -1, 1:-1] = N.EMPTY
    gen = torch.Generator(device="cuda:0").manual_seed(SEED)
    with N.Solver() as s:
        s.commit(g)
        nu, nv, nw, npp = s.field_sizes()
        u0 = torch.randn(nu, dtype=torch.float64, device="cuda:0", generator=gen)
        v0 = torch.randn(nv, dtype=torch.float64, device="cuda:0", generator=gen)
        w0 = torch.randn(nw, dtype=torch.float64, device="cuda:0", generator=gen)
        u, v, w = u0.clone(), v0.clone(), w0.clone()
        p = torch.zeros(npp, dtype=torch.float64, device="cuda:0")
        b = torch.zeros(s.n, dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        rho, dt = 1.0, 1e-3
        cells = float(npp)
        faces = float(nu + nv + nw)
        rhs_bytes = 8.0 * faces + cells + 8.0 * s.n
        proj_bytes = 16.0 * faces + 9.0 * cells

        def timed(fn, reps):
            for _ in range(max(1, args.warmup)):
                fn()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            return (time.perf_counter() - t0) / reps

        reps = max(20, 5 * args.steps)
        t_rhs = timed(lambda: s.rhs_device(u.data_ptr(), v.data_ptr(), w.data_ptr(), p.data_ptr(), 0, rho, dt,
                                           b.data_ptr()), reps)
        t_proj = timed(lambda: s.project_device(u.data_ptr(), v.data_ptr(), w.data_ptr(), p.data_ptr(), rho, dt), reps)

        def step():
            u.copy_(u0)
            v.copy_(v0)
            w.copy_(w0)
            torch.cuda.synchronize()
            res, rc = s.pressure_step_device(u.data_ptr(), v.data_ptr(), w.data_ptr(), p.data_ptr(), 0, rho, dt)
            if rc != 0 or not res.relative_residual < 1e-8:
                raise SystemExit(f"pressure step did not converge: {res.relative_residual}")
            s.projection_step_device(u.data_ptr(), v.data_ptr(), w.data_ptr(), p.data_ptr(), 0, rho, dt)
            return res

        res = step()
        t_step = timed(step, args.steps)
        # what the projection left: rho/dt div(u) V with p = 0, against the divergence it started from
        z = torch.zeros_like(p)
        s.rhs_device(u.data_ptr(), v.data_ptr(), w.data_ptr(), z.data_ptr(), 0, rho, dt, b.data_ptr())
        after = float(torch.linalg.norm(b))
        s.rhs_device(u0.data_ptr(), v0.data_ptr(), w0.data_ptr(), z.data_ptr(), 0, rho, dt, b.data_ptr())
        before = float(torch.linalg.norm(b))
        # the host-array variants (PCIe in and out)
        hu, hv, hw, hp = (t.cpu().numpy() for t in (u0, v0, w0, p))
        t_rhs_host = timed(lambda: s.rhs(hu, hv, hw, hp, rho, dt), 3)
        t_proj_host = timed(lambda: s.project(hu, hv, hw, hp, rho, dt), 3)
    cpu = None
    if not args.no_cpu:
        import oracle as O
        og = O.Grid(3, g.types, g.wx, g.wy, g.wz, g.dh, False)
        sysm = O.assemble(og)
        t0 = time.perf_counter()
        O.rhs(og, sysm, hu, hv, hw, hp, rho, dt)
        t_cr = time.perf_counter() - t0
        t0 = time.perf_counter()
        O.project(og, hu, hv, hw, hp, rho, dt)
        t_cp = time.perf_counter() - t0
        cpu = {"rhs_ms": 1e3 * t_cr, "project_ms": 1e3 * t_cp, "cores": 1, "kind": "port",
               "sample": f"one pass of the oracle's C loops (oracle_rhs, oracle_project) over the {n}^3 fields"}
    return {"metric": "pressure step (RHS + solve + scatter + projection) per second at 256^3", "value": 1.0 / t_step,
            "unit": "steps/s", "ms_per_step": 1e3 * t_step,
            "rhs_ms": 1e3 * t_rhs, "rhs_gbps": rhs_bytes / t_rhs / 1e9, "rhs_bytes": rhs_bytes,
            "project_ms": 1e3 * t_proj, "project_gbps": proj_bytes / t_proj / 1e9, "project_bytes": proj_bytes,
            "rhs_host_ms": 1e3 * t_rhs_host, "project_host_ms": 1e3 * t_proj_host,
            "iterations": int(res.iterations), "relative_residual": float(res.relative_residual),
            "divergence_before": before, "divergence_after": after, "cpu_baseline": cpu,
            "config": {"workload": f"{n}^3 uniform box, SOLID walls, EMPTY (p = 0) outlet ring at x-max; N(0,1) "
                                   f"MAC velocity (torch seed {SEED}); rho = 1, dt = 1e-3; Chorin step: RHS, MGPCG to "
                                   f"1e-8, scatter + EnforcePressureBC, ProjectPressure",
                       "unknowns": int(s.n)}}


def layout_leg(args):
    """FLIP's layout change at 128^3: a free-surface box whose liquid moves every step. Times a full commit of each
    new layout against nppe_update_layout (host types) and nppe_update_layout_device (types in HBM), and the solve
    that follows (to 1e-8)."""
    import torch
    import nereus_cfd_amd as N
    import nereus_cfd_amd.workloads as W
    n = 128 if args.size == 256 else args.size
    shape = (n, n, n)
    grids = [W.flip_free_surface(shape, np.random.default_rng(SEED + q)) for q in range(args.steps + 1)]
    dev = [torch.from_numpy(np.ascontiguousarray(g.types)).to("cuda:0") for g in grids]
    torch.cuda.synchronize()
    t_commit, t_host, t_dev, t_same, t_solve, codes = [], [], [], [], [], []
    with N.Solver() as fresh, N.Solver() as s, N.Solver() as d:
        s.commit(grids[0])
        d.commit(grids[0])
        for q in range(1, len(grids)):
            t0 = time.perf_counter()
            fresh.commit(grids[q])
            t_commit.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            codes.append(s.update_layout(grids[q].types))
            t_host.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            d.update_layout_device(dev[q].data_ptr(), grids[q].types)
            t_dev.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            d.update_layout_device(dev[q].data_ptr(), grids[q].types)
            t_same.append(time.perf_counter() - t0)
            b = W.random_rhs(grids[q], np.random.default_rng(q), zero_mean=False)
            tb = torch.from_numpy(b).cuda()
            tx = torch.zeros_like(tb)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res, rc = d.solve_device(tb.data_ptr(), tx.data_ptr())
            t_solve.append(time.perf_counter() - t0)
            if rc != 0 or not res.relative_residual < 1e-8:
                raise SystemExit(f"solve after the layout update did not converge: {res.relative_residual}")
    med = lambda v: 1e3 * float(np.median(v))
    return {"metric": "FLIP layout change + solve per second at 128^3", "value": 1.0 / (np.median(t_dev) +
                                                                                       np.median(t_solve)),
            "unit": "steps/s", "ms_per_step": med(t_dev) + med(t_solve),
            "commit_ms": med(t_commit), "update_host_ms": med(t_host), "update_device_ms": med(t_dev),
            "unchanged_check_ms": med(t_same), "solve_ms": med(t_solve), "update_codes": codes,
            "config": {"workload": f"{n}^3 box, SOLID walls, liquid = dam-break block (30-50 % of the volume) + 0-200 "
                                   f"spheres r = 2-6 cells, EMPTY elsewhere (a new random layout each step, seeds "
                                   f"{SEED}+q); MGPCG to 1e-8 after each update",
                       "steps": args.steps}}


def batched_leg(args):
    """32 FLIP-like free-surface systems of sizes U{64..160}^3, solved together to 1e-8: nppe_batch_solve with device
    vectors, which shares every launch of the inner iteration between the systems (one per V-cycle place and kernel,
    replayed as one graph)."""
    import torch
    import nereus_cfd_amd as N
    import nereus_cfd_amd.workloads as W
    rng = np.random.default_rng(SEED)
    grids = [W.flip_free_surface(shape, rng) for shape in W.batch_shapes(32, rng)]
    bs = [W.random_rhs(g, rng, zero_mean=False) for g in grids]
    solvers = [N.Solver(tolerance=1e-8) for _ in grids]
    t0 = time.perf_counter()
    for h, g in zip(solvers, grids):
        h.commit(g)
    setup = time.perf_counter() - t0
    tbs = [torch.from_numpy(b).cuda() for b in bs]
    txs = [torch.zeros_like(t) for t in tbs]
    torch.cuda.synchronize()
    bp, xp = [t.data_ptr() for t in tbs], [t.data_ptr() for t in txs]
    # the first call records the systems' iterations, merges them into shared launches and times the shared
    # sweeps' tiles (the batch's plan), then solves
    tp = time.perf_counter()
    N.batch_solve_device(solvers, bp, xp)
    first_call = time.perf_counter() - tp
    for _ in range(max(0, args.warmup - 1)):
        N.batch_solve_device(solvers, bp, xp)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, rc = N.batch_solve_device(solvers, bp, xp)
        if rc != 0:
            raise SystemExit(f"batched solve failed: {rc}")
    el = time.perf_counter() - t0
    # the same systems one after another, each on its own handle's stream, for comparison (one untimed pass first:
    # a handle solving alone captures its own iteration graphs at its first solve)
    for h, b, x in zip(solvers, bp, xp):
        h.solve_device(b, x)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for h, b, x in zip(solvers, bp, xp):
        h.solve_device(b, x)
    seq = time.perf_counter() - t1
    by = sum(r.algorithmic_bytes for r in res)
    for h in solvers:
        h.close()
    cells = sum(int(np.prod(g.types.shape)) for g in grids)
    return {"metric": "batched free-surface MGPCG solves/s (32 systems, 64^3-160^3, to 1e-8)",
            "value": 32 * args.steps / el, "unit": "solves/s", "ms_per_step": 1e3 * el / args.steps,
            "gbps": by / (el / args.steps) / 1e9, "sequential_ms_per_batch": 1e3 * seq,
            "speedup_vs_sequential": seq / (el / args.steps), "first_call_ms": 1e3 * first_call,
            "max_iterations": int(max(r.iterations for r in res)), "worst_relative_residual":
                float(max(r.relative_residual for r in res)), "setup_s": setup,
            "config": {"workload": f"32 FLIP-like 3D grids, sizes U{{64..160}} per axis, liquid block 30-50% + "
                                   f"0-200 spheres r 2-6, EMPTY air (p=0), SOLID walls, N(0,1) RHS seed {SEED}",
                       "cells": cells}}


# The serving stream's resident grid set: (dim, cells per axis, boundary condition), spanning the stated mix.
SERVING_GRIDS = [(2, 256, "outlet"), (2, 512, "neumann"), (2, 1024, "freesurface"), (2, 2048, "outlet"),
                 (3, 64, "freesurface"), (3, 96, "neumann"), (3, 128, "outlet"), (3, 192, "freesurface"),
                 (3, 256, "neumann")]


def serving_leg(args):
    """The serving stream on a fixed set of resident grids (2D 256^2-2048^2, 3D 64^3-256^3; Neumann / p=0 outlet /
    free surface), each request a seeded right-hand side and a tolerance 10^U(-10,-6) on a uniformly drawn grid:
    1. first use: one request per grid, its latency including the commit (reported apart, as commit latency);
    2. capacity: every request submitted at once (closed loop), solves/s;
    3. serving: Poisson arrivals at 70% of that capacity (open loop), solves/s and p50/p99 latency from submit to
       completion; the FP64 CG server on the same stream at the same rate, beside it."""
    import threading
    import nereus_cfd_amd as N
    import nereus_cfd_amd.workloads as W
    rng = np.random.default_rng(SEED)
    grids = [W.serving_grid(d, n, bc, seed=k) for k, (d, n, bc) in enumerate(SERVING_GRIDS)]
    pool = [[W.random_rhs(g, np.random.default_rng(SEED + 7 * k + v), zero_mean=(SERVING_GRIDS[k][2] == "neumann"))
             for v in range(2)] for k, g in enumerate(grids)]
    nreq = args.requests
    reqs = [(int(rng.integers(0, len(grids))), float(10.0 ** rng.uniform(-10.0, -6.0)), int(rng.integers(0, 2)))
            for _ in range(nreq)]
    gaps = rng.exponential(1.0, size=nreq)   # unit-rate gaps, scaled to the serving rate below

    def stream(srv, rate):
        """Submits reqs with Poisson gaps at `rate` (inf: all at once); returns (latencies, wall, all converged)."""
        tickets = [None] * nreq
        arrivals = np.cumsum(gaps / rate) if np.isfinite(rate) else np.zeros(nreq)
        t0 = time.perf_counter()

        def submitter():
            for q, (k, tol, v) in enumerate(reqs):
                delay = arrivals[q] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                tickets[q] = srv.submit(k, pool[k][v], tol)

        th = threading.Thread(target=submitter)
        th.start()
        lats, ok = [], True
        for q in range(nreq):
            while tickets[q] is None:
                time.sleep(0.0002)
            x, res, lat, rc = srv.wait(tickets[q])
            lats.append(lat)
            ok = ok and rc == 0 and res.relative_residual <= reqs[q][1] * 1.0000001
        th.join()
        return np.array(lats), time.perf_counter() - t0, ok

    def server(method):
        srv = N.Server(slots=args.slots, hbm_budget=int(args.budget_gb * 2 ** 30), method=method, max_iters=100000)
        first = []
        for k, g in enumerate(grids):
            srv.add_grid(k, g)
            t = srv.submit(k, pool[k][0], 1e-8)
            x, res, lat, rc = srv.wait(t)
            first.append(lat)
        return srv, first

    srv, first = server(N.METHOD_MGPCG)
    lat_c, wall_c, ok_c = stream(srv, float("inf"))
    capacity = nreq / wall_c
    rate = args.rate if args.rate > 0 else 0.7 * capacity
    lat, wall, ok = stream(srv, rate)
    peak = srv.memory()[1]
    srv.close()
    log(f"[bench] serving: capacity {capacity:.1f} solves/s; at {rate:.1f}/s p50 {np.percentile(lat, 50):.1f} ms "
        f"p99 {np.percentile(lat, 99):.1f} ms")
    cg = None
    if not args.no_cg64:
        srv, first_cg = server(N.METHOD_CG64)
        lat_g, wall_g, ok_g = stream(srv, rate)
        srv.close()
        cg = {"solves_per_s": nreq / wall_g, "p50_ms": float(np.percentile(lat_g, 50)),
              "p99_ms": float(np.percentile(lat_g, 99)), "all_converged": ok_g, "first_request_ms": first_cg}
    return {"metric": "serving solves/s and p50/p99 latency at a fixed Poisson arrival rate",
            "value": nreq / wall, "unit": "solves/s", "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "arrival_rate_per_s": rate, "capacity_solves_per_s": capacity,
            "capacity_p50_ms": float(np.percentile(lat_c, 50)), "all_converged": ok and ok_c,
            "first_request_ms": {f"{d}D {n} {bc}": round(v, 1) for (d, n, bc), v in zip(SERVING_GRIDS, first)},
            "peak_hbm_gb": peak / 2 ** 30, "fp64_cg": cg,
            "config": {"workload": f"{nreq} requests on {len(grids)} resident grids (2D 256^2-2048^2, 3D 64^3-256^3; "
                                   f"Neumann / p=0 outlet / free surface), tolerance 10^U(-10,-6); Poisson arrivals "
                                   f"at {rate:.1f}/s = {rate / capacity:.0%} of the measured capacity; {args.slots} "
                                   f"slots, HBM budget {args.budget_gb} GiB; first use (commit) reported apart",
                       "seed": SEED}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cg64", action="store_true")
    ap.add_argument("--config", default="headline", choices=["headline", "channel", "batched", "serving", "copy", "step", "layout"])
    ap.add_argument("--requests", type=int, default=64)
    ap.add_argument("--rate", type=float, default=0.0, help="serving arrival rate (0: 70%% of the measured capacity)")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    if args.config != "headline":
        leg = {"channel": channel_leg, "batched": batched_leg, "serving": serving_leg, "copy": copy_leg,
               "step": step_leg, "layout": layout_leg}[args.config]
        out = leg(args)
        out.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
                    "dtype": "fp32+fp64", "data": "synthetic", "vs_baseline": None})
        print(json.dumps(out))
        return
    res = gpu_leg(args)
    cpu = None if args.no_cpu else cpu_leg(args)
    if cpu and res.get("fp64_cg"):
        # Eigen's CG needs the iterations the same Jacobi-CG takes on the GPU (Eigen's semantics, pinned to within
        # one on the golden systems): its projected full solve
        its = res["fp64_cg"]["iterations"]
        cpu["eigen_cg_projected_s"] = cpu["eigen_cg_s_per_iter"] * its
        cpu["eigen_cg_projected_iterations"] = its
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
                               f"MGPCG: FP32 V(2,2) red-black Gauss-Seidel multigrid preconditioner in an FP64 CG "
                               f"recursion, to 1e-8 true FP64 relative residual",
                   "cells": args.size ** 3, "tolerance": 1e-8},
        "roofline": res.pop("roofline"),
        "cpu_baseline": cpu,
    }
    line.update(res)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
