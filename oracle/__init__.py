"""CPU oracle for the pressure Poisson path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this package, and only as
the checker. The product library (``nereus_cfd_amd``) never imports it.

* ``liboracle.so`` (``oracle.c``): a plain-C restatement of the reference's spacing tables (MAC.cpp:19-175),
  matrix assembly (PressureSolver3D.cpp:128-178, PressureSolver2D.cpp:100-152) and Eigen 3.3.7 Jacobi-PCG
  (ConjugateGradient.h:28-91), wrapped here with numpy.
* ``_ref/ref_harness``: the reference's own sources compiled in place (``make -C oracle ref``); ``refcase``
  drives it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

FLUID, SOLID, EMPTY, INFLOW, PRESSURE, OUTFLOW = 0, 1, 2, 3, 4, 5
REFINE_HARD, REFINE_RAMP, REFINE_SYMMETRIC, REFINE_RAMP_BACK = 0, 1, 2, 3


class _Ref(C.Structure):
    _fields_ = [("start", C.c_double), ("end", C.c_double), ("factor", C.c_int32), ("mode", C.c_int32)]


class _Grid(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32),
        ("nonUniform", C.c_int32), ("dh", C.c_double),
        ("types", C.c_void_p), ("wx", C.c_void_p), ("wy", C.c_void_p), ("wz", C.c_void_p),
    ]


_lib = None


def build():
    """Compiles liboracle.so (gcc)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.oracle_build_axis.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, P, C.c_int, C.c_int, P, C.c_int]
        L.oracle_build_axis.restype = C.c_int
        L.oracle_gap_table.argtypes = [P, C.c_int, P]
        L.oracle_count_unknowns.argtypes = [P]
        L.oracle_count_unknowns.restype = C.c_int64
        L.oracle_assemble.argtypes = [P, P, P, P, P, P]
        L.oracle_assemble.restype = C.c_int64
        L.oracle_spmv.argtypes = [C.c_int32, P, P, P, P, P]
        L.oracle_cg.argtypes = [C.c_int32, P, P, P, P, P, C.c_double, C.c_int, C.c_int, P]
        L.oracle_cg.restype = C.c_int
        L.oracle_solve_pressure.argtypes = [C.c_int32, P, P, P, P, P, P, C.c_double, C.c_int, C.c_double]
        L.oracle_solve_pressure.restype = C.c_int
        L.oracle_subtract_mean.argtypes = [P, C.c_int64]
        L.oracle_relres.argtypes = [C.c_int32, P, P, P, P, P]
        L.oracle_relres.restype = C.c_double
        L.oracle_rhs.argtypes = [P, P, C.c_int64, P, P, P, P, P, C.c_double, C.c_double, P]
        L.oracle_project.argtypes = [P, P, P, P, P, C.c_double, C.c_double]
        L.oracle_enforce_pressure_bc.argtypes = [P, P, P]
        L.oracle_scatter_pressure.argtypes = [P, P, C.c_int64, P, P, P, C.c_int, P]
        L.oracle_accumulate_increment.argtypes = [P, P, C.c_int64, P, P, C.c_int, P]
        L.oracle_level_inflow.argtypes = [P, P, C.c_int64, P, P, P, P, P]
        L.oracle_level_inflow.restype = C.c_int
        L.oracle_warm_start_guess.argtypes = [P, P, C.c_int64, P, C.c_int, P]
        L.oracle_warm_start_guess.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


@dataclass
class Refinement:
    start: float
    end: float
    factor: int
    mode: int = REFINE_HARD


def build_axis(a0, af, n_base, dh, refs=(), non_uniform=False):
    """Widths of one axis including the ring (MAC::BuildAxis + MAC::AddRing)."""
    arr = (_Ref * max(1, len(refs)))(*[_Ref(r.start, r.end, r.factor, r.mode) for r in refs])
    cap = 4 * n_base * max([1] + [int(r.factor) for r in refs]) + 16
    out = np.zeros(cap, np.float64)
    n = lib().oracle_build_axis(a0, af, n_base, dh, C.cast(arr, C.c_void_p), len(refs), int(non_uniform),
                                _p(out), cap)
    if n < 0:
        raise ValueError("axis too long")
    return out[:n].copy()


def gap_table(w):
    w = np.ascontiguousarray(w, np.float64)
    g = np.zeros(len(w) + 1, np.float64)
    lib().oracle_gap_table(_p(w), len(w), _p(g))
    return g


@dataclass
class Grid:
    """A MAC grid as the pressure path sees it: dims incl. the ring, cell types and width tables."""
    dim: int
    types: np.ndarray          # uint8, shape (Ny, Nx, Nz) (Nz = 1 in 2D)
    wx: np.ndarray
    wy: np.ndarray
    wz: np.ndarray
    dh: float
    non_uniform: bool = False

    @property
    def shape(self):
        return self.types.shape

    def _c(self):
        Ny, Nx, Nz = self.types.shape
        self._keep = [np.ascontiguousarray(self.types, np.uint8), np.ascontiguousarray(self.wx, np.float64),
                      np.ascontiguousarray(self.wy, np.float64), np.ascontiguousarray(self.wz, np.float64)]
        t, wx, wy, wz = self._keep
        return _Grid(self.dim, Nx, Ny, Nz, int(self.non_uniform), self.dh, t.ctypes.data, wx.ctypes.data,
                     wy.ctypes.data, wz.ctypes.data)


@dataclass
class System:
    n: int
    order: np.ndarray
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    singular: bool
    extra: dict = field(default_factory=dict)

    def matvec(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.n, np.float64)
        lib().oracle_spmv(self.n, _p(self.rowptr), _p(self.col), _p(self.val), _p(x), _p(y))
        return y

    def relres(self, b, x):
        b = np.ascontiguousarray(b, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        return lib().oracle_relres(self.n, _p(self.rowptr), _p(self.col), _p(self.val), _p(b), _p(x))

    def cg(self, b, tol, max_iters=100000, x0=None, precond="diagonal"):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n, np.float64) if x0 is None else np.array(x0, np.float64)
        err = np.zeros(1, np.float64)
        it = lib().oracle_cg(self.n, _p(self.rowptr), _p(self.col), _p(self.val), _p(b), _p(x), tol, max_iters,
                             0 if precond == "diagonal" else 1, _p(err))
        return x, it, float(err[0])

    def solve_pressure(self, b, tol, max_iters=100000, guess=None, reference_norm=0.0):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.n, np.float64)
        g = None if guess is None else np.ascontiguousarray(guess, np.float64)
        it = lib().oracle_solve_pressure(self.n, _p(self.rowptr), _p(self.col), _p(self.val), _p(b), _p(x), _p(g),
                                         tol, max_iters, reference_norm)
        return x, it

    def dense(self):
        A = np.zeros((self.n, self.n))
        for r in range(self.n):
            for q in range(self.rowptr[r], self.rowptr[r + 1]):
                A[r, self.col[q]] = self.val[q]
        return A


def assemble(grid: Grid) -> System:
    g = grid._c()
    L = lib()
    n = int(L.oracle_count_unknowns(C.byref(g)))
    nq = 7 if grid.dim == 3 else 5
    order = np.zeros(max(n, 1), np.int32)
    rowptr = np.zeros(n + 1, np.int32)
    col = np.zeros(max(1, nq * n), np.int32)
    val = np.zeros(max(1, nq * n), np.float64)
    sing = np.zeros(1, np.int32)
    nnz = int(L.oracle_assemble(C.byref(g), _p(order), _p(rowptr), _p(col), _p(val), _p(sing)))
    return System(n, order[:n].copy(), rowptr, col[:nnz].copy(), val[:nnz].copy(), bool(sing[0]))


def subtract_mean(v):
    v = np.ascontiguousarray(v, np.float64).copy()
    lib().oracle_subtract_mean(_p(v), len(v))
    return v


def compatibility(b):
    """The RHS compatibility step: the mean removed twice (PressureSolver3D.cpp:300-307)."""
    b = np.array(b, np.float64)
    if len(b) == 0:
        return b
    return subtract_mean(subtract_mean(b))


def field_sizes(grid: Grid):
    """Lengths of the MAC arrays u, v, w, p (MAC.h:235-238; 2D :369-371): w is empty in 2D."""
    Ny, Nx, Nz = grid.types.shape
    if grid.dim == 2:
        return Ny * (Nx + 1), (Ny + 1) * Nx, 0, Ny * Nx
    return Ny * (Nx + 1) * Nz, (Ny + 1) * Nx * Nz, Ny * Nx * (Nz + 1), Ny * Nx * Nz


def rhs(grid: Grid, system: System, u, v, w, p, rho, dt, p_old=None):
    """b = rho/dt * div(u) * V - scale * DirichletSum in IDP order (oracle_rhs)."""
    g = grid._c()
    arrs = [np.ascontiguousarray(a, np.float64) if a is not None else None for a in (u, v, w, p, p_old)]
    if arrs[2] is None:
        arrs[2] = np.zeros(1)
    b = np.zeros(max(1, system.n), np.float64)
    lib().oracle_rhs(C.byref(g), _p(system.order), system.n, *[_p(a) for a in arrs], rho, dt, _p(b))
    return b[:system.n]


def project(grid: Grid, u, v, w, p, rho, dt):
    """(u, v, w) after u -= dt/rho * grad p (oracle_project); returns copies."""
    g = grid._c()
    u = np.array(u, np.float64)
    v = np.array(v, np.float64)
    w = np.array(w if w is not None else np.zeros(1), np.float64)
    p = np.ascontiguousarray(p, np.float64)
    lib().oracle_project(C.byref(g), _p(u), _p(v), _p(w), _p(p), rho, dt)
    return u, v, (w if grid.dim == 3 else None)


def _opt64(a):
    return None if a is None else np.ascontiguousarray(a, np.float64)


def enforce_pressure_bc(grid: Grid, p, prescribed=None):
    """p (a copy) after MAC::EnforcePressureBC (oracle_enforce_pressure_bc); prescribed: box-sized values for the
    PRESSURE ring cells, or None to keep theirs."""
    g = grid._c()
    p = np.array(p, np.float64)
    pr = _opt64(prescribed)
    lib().oracle_enforce_pressure_bc(C.byref(g), _p(p), _p(pr))
    return p


def scatter_pressure(grid: Grid, system: System, x, p, p_old=None, level=False, prescribed=None):
    """p (a copy) after SolveScheme's tail with x as the solve's answer (oracle_scatter_pressure)."""
    g = grid._c()
    p = np.array(p, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    po, pr = _opt64(p_old), _opt64(prescribed)
    lib().oracle_scatter_pressure(C.byref(g), _p(system.order), system.n, _p(x), _p(p), _p(po), int(bool(level)),
                                  _p(pr))
    return p


def accumulate_increment(grid: Grid, system: System, p_old, p, level=False, prescribed=None):
    """p (a copy) after AccumulatePressureIncrement (oracle_accumulate_increment)."""
    g = grid._c()
    p = np.array(p, np.float64)
    po, pr = np.ascontiguousarray(p_old, np.float64), _opt64(prescribed)
    lib().oracle_accumulate_increment(C.byref(g), _p(system.order), system.n, _p(po), _p(p), int(bool(level)),
                                      _p(pr))
    return p


def level_pressure(grid: Grid, system: System, x, level=None, centres=None):
    """LevelPressure (PressureSolver3D.cpp:181-211) on a copy of x with the compatibility condition in force: the
    mean removed, then (level: the scene's value at each ring cell, NaN where none; centres: (cx, cy, cz) of
    MAC::CentrePos) the inflow shift (oracle_level_inflow). Returns (x, contributing ring cells)."""
    x = np.array(x, np.float64)
    x -= x.mean() if x.size else 0.0
    cnt = 0
    if level is not None:
        g = grid._c()
        lv = np.ascontiguousarray(level, np.float64)
        cx, cy, cz = (np.ascontiguousarray(c, np.float64) for c in centres)
        cnt = lib().oracle_level_inflow(C.byref(g), _p(system.order), system.n, _p(x), _p(lv), _p(cx), _p(cy), _p(cz))
    return x, cnt


def warm_start_guess(grid: Grid, system: System, p_prev, level):
    """BuildWarmStartGuess once enabled (oracle_warm_start_guess): (built, x0)."""
    g = grid._c()
    x0 = np.zeros(max(1, system.n), np.float64)
    pp = np.ascontiguousarray(p_prev, np.float64)
    built = lib().oracle_warm_start_guess(C.byref(g), _p(system.order), system.n, _p(pp), int(bool(level)), _p(x0))
    return bool(built), x0[: system.n]
