"""Drives ``_ref/ref_harness`` (the reference's own pressure path, compiled from its sources by ``make -C oracle
ref``). Test infrastructure only: used to write golden fixtures and for bench.py's CPU-baseline leg.

Solver codes follow the reference's PRESSURE block keys (PressureSettings.h:44-45): eigen_solver BICGSTAB=0, CG=1,
AMG=2 (AMGCL CG + smoothed aggregation, SPAI0); eigen_preconditioner DIAGONAL=0, IDENTITY=1.
"""
from __future__ import annotations

import os
import struct
import subprocess
import tempfile
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import HERE, Refinement

HARNESS = os.path.join(HERE, "_ref", "ref_harness")
BICGSTAB, CG, AMG = 0, 1, 2
DIAGONAL, IDENTITY = 0, 1


def available() -> bool:
    return os.path.exists(HARNESS)


@dataclass
class SolveSpec:
    solver: int = CG
    precond: int = DIAGONAL
    tol: float = 1e-12
    max_iters: int = 2000
    ref_norm: float = 0.0
    repeats: int = 1
    guess: Optional[np.ndarray] = None   # SolvePressureSystem's guess (n values), or none


@dataclass
class RefCase:
    dim: int
    grid_size: int
    domain: Sequence[float] = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    non_uniform: bool = False
    refine: Sequence[Sequence[Refinement]] = ((), (), ())
    types: Optional[np.ndarray] = None        # uint8 (Ny, Nx, Nz)
    compat: bool = False
    solves: List[SolveSpec] = field(default_factory=list)
    rhs: Optional[np.ndarray] = None
    # velocity and pressure fields in MAC layout: {"rho", "dt", "u", "v", "w", "p", "p_old" (optional)}; the
    # harness then returns the right-hand side built from them, the increment reference norm (with p_old) and the
    # projected fields; with "t" (time) and "x_tail" (n values standing for a solve's answer) also p after
    # EnforcePressureBC, after SolveScheme's tail and (with p_old) after AccumulatePressureIncrement
    fields: Optional[dict] = None
    # LevelPressure with a scene's inlet levels and BuildWarmStartGuess: {"scene": text (inlet faces with p=),
    # "t": time, "x": n values to level, "p_prev": MAC pressure for the warm start (optional)}; the harness returns
    # the scene's level at every ring cell, the centre tables, the levelled x and the warm-start guess
    extras: Optional[dict] = None
    # timing runs of large systems: leave the unknown order and the matrix (and the solutions) out of the output
    write_matrix: bool = True
    write_solutions: bool = True

    def encode(self) -> bytes:
        out = [b"NPPECASE", struct.pack("<ii", self.dim, self.grid_size), struct.pack("<6d", *self.domain),
               struct.pack("<i", int(self.non_uniform)),
               struct.pack("<3i", *[len(r) for r in self.refine])]
        for axis in self.refine:
            for r in axis:
                out.append(struct.pack("<ddii", r.start, r.end, r.factor, r.mode))
        if self.types is None:
            out.append(struct.pack("<i", 0))
            return b"".join(out)
        Ny, Nx, Nz = self.types.shape
        out.append(struct.pack("<iiii", 1, Nx, Ny, Nz))
        out.append(np.ascontiguousarray(self.types, np.uint8).tobytes())
        out.append(struct.pack("<i", int(self.compat)))
        out.append(struct.pack("<i", len(self.solves)))
        for s in self.solves:
            out.append(struct.pack("<iidid", s.solver, s.precond, s.tol, s.max_iters, s.ref_norm))
            out.append(struct.pack("<ii", s.repeats, 0 if s.guess is None else 1))
        if self.solves:
            out.append(np.ascontiguousarray(self.rhs, np.float64).tobytes())
        for s in self.solves:
            if s.guess is not None:
                out.append(np.ascontiguousarray(s.guess, np.float64).tobytes())
        f = self.fields
        out.append(struct.pack("<i", 1 if f else 0))
        if f:
            out.append(struct.pack("<ddi", f["rho"], f["dt"], 1 if f.get("p_old") is not None else 0))
            for key in ("u", "v") + (("w",) if self.dim == 3 else ()) + ("p",):
                out.append(np.ascontiguousarray(f[key], np.float64).tobytes())
            if f.get("p_old") is not None:
                out.append(np.ascontiguousarray(f["p_old"], np.float64).tobytes())
            out.append(struct.pack("<d", f.get("t", 0.0)))
            out.append(np.ascontiguousarray(f["x_tail"], np.float64).tobytes())
        e = self.extras
        out.append(struct.pack("<i", 1 if e else 0))
        if e:
            text = e.get("scene", "").encode()
            out.append(struct.pack("<i", len(text)))
            out.append(text)
            out.append(struct.pack("<d", e.get("t", 0.0)))
            out.append(np.ascontiguousarray(e["x"], np.float64).tobytes())
            out.append(struct.pack("<i", 1 if e.get("p_prev") is not None else 0))
            if e.get("p_prev") is not None:
                out.append(np.ascontiguousarray(e["p_prev"], np.float64).tobytes())
        out.append(struct.pack("<i", (0 if self.write_matrix else 1) | (0 if self.write_solutions else 2)))
        return b"".join(out)


class _Buf:
    def __init__(self, data: bytes):
        self.d, self.o = data, 0

    def take(self, fmt):
        v = struct.unpack_from(fmt, self.d, self.o)
        self.o += struct.calcsize(fmt)
        return v

    def arr(self, dtype, n):
        a = np.frombuffer(self.d, dtype, n, self.o).copy()
        self.o += a.nbytes
        return a


def run(case: RefCase, timeout: float = 3600.0, threads: Optional[int] = None) -> dict:
    """Runs the harness on one case and returns its tables, system and solves."""
    if not available():
        raise FileNotFoundError(HARNESS + " is not built (make -C oracle ref)")
    env = dict(os.environ)
    if threads is not None:
        env["OMP_NUM_THREADS"] = str(threads)
    with tempfile.TemporaryDirectory() as td:
        cin, cout = os.path.join(td, "case.bin"), os.path.join(td, "out.bin")
        with open(cin, "wb") as f:
            f.write(case.encode())
        subprocess.run([HARNESS, cin, cout], check=True, cwd=td, env=env, timeout=timeout,
                       stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
        with open(cout, "rb") as f:
            data = f.read()
    b = _Buf(data)
    assert b.take("8s")[0] == b"NPPEOUT1"
    Nx, Ny, Nz, nu = b.take("<iiii")
    dh = b.take("<d")[0]
    res = {"Nx": Nx, "Ny": Ny, "Nz": Nz, "non_uniform": bool(nu), "dh": dh}
    axes = [("x", Nx), ("y", Ny)] + ([("z", Nz)] if case.dim == 3 else [])
    for name, N in axes:
        res["w" + name] = b.arr(np.float64, N)
        res["gap" + name] = b.arr(np.float64, N + 1)
    if case.dim == 2:
        res["wz"] = np.array([dh])
    if case.types is None:
        return res
    n, nnz, singular = b.take("<iii")
    res.update(n=n, nnz=nnz, singular=bool(singular))
    if case.write_matrix:
        res["order"] = b.arr(np.int32, n)
        res["rowptr"] = b.arr(np.int32, n + 1)
        res["col"] = b.arr(np.int32, nnz)
        res["val"] = b.arr(np.float64, nnz)
    nsolves = b.take("<i")[0]
    res["solves"] = []
    for _ in range(nsolves):
        iters, nt = b.take("<ii")
        times = b.arr(np.float64, nt)
        x = b.arr(np.float64, n) if case.write_solutions else None
        res["solves"].append({"iters": iters, "times": times, "x": x})
    res["threads"] = b.take("<i")[0]
    if case.fields:
        Nzz = Nz if case.dim == 3 else 1
        res["rhs_built"] = b.arr(np.float64, n)
        res["ref_norm"] = b.take("<d")[0]
        res["u"] = b.arr(np.float64, Ny * (Nx + 1) * Nzz)
        res["v"] = b.arr(np.float64, (Ny + 1) * Nx * Nzz)
        if case.dim == 3:
            res["w"] = b.arr(np.float64, Ny * Nx * (Nz + 1))
        cells = Ny * Nx * Nzz
        res["p_bc"] = b.arr(np.float64, cells)
        res["p_tail"] = b.arr(np.float64, cells)
        if case.fields.get("p_old") is not None:
            res["p_acc"] = b.arr(np.float64, cells)
    if case.extras:
        Nzz = Nz if case.dim == 3 else 1
        res["scene_loaded"] = bool(b.take("<i")[0])
        res["inflow_level"] = b.arr(np.float64, Ny * Nx * Nzz)
        res["cx"] = b.arr(np.float64, Nx)
        res["cy"] = b.arr(np.float64, Ny)
        res["cz"] = b.arr(np.float64, Nz) if case.dim == 3 else np.zeros(1)
        res["x_levelled"] = b.arr(np.float64, n)
        if b.take("<i")[0]:
            res["warm_built"] = bool(b.take("<i")[0])
            res["warm_guess"] = b.arr(np.float64, n)
    return res
