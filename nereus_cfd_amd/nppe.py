"""ctypes binding of libnppe.so, the C ABI declared in include/nppe.h, include/nppe_profile.h and
include/nppe_synth.h.

The library is the product; this module only marshals numpy arrays (host) or device pointers (e.g. from torch) to
it. Loading fails loudly when libnppe.so is missing: there is no fallback solver.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnppe.so")

OK, NOT_CONVERGED = 0, 1
ERR_ARGUMENT, ERR_DEVICE, ERR_NO_GRID, ERR_SIZE, ERR_NONFINITE, ERR_BUDGET, ERR_MISMATCH, ERR_INTERNAL = (
    -1, -2, -3, -4, -5, -6, -7, -8)
METHOD_MGPCG, METHOD_CG64 = 0, 1
KRYLOV_REFINE, KRYLOV_FP64 = 0, 1
FLUID, SOLID, EMPTY, INFLOW, PRESSURE, OUTFLOW = 0, 1, 2, 3, 4, 5

KERNEL_KINDS = [
    "smooth", "restrict", "prolong", "coarse", "cg_dot", "cg_direction", "cg_spmv", "cg_update", "residual64",
    "scatter", "gather", "mean", "cg64_init", "cg64_spmv", "cg64_update", "cg64_direction", "smooth0",
    "smooth_restrict",
]

# Every symbol the headers in include/ declare.
EXPORTS = [
    "nppe_default_options", "nppe_create", "nppe_destroy", "nppe_last_error", "nppe_commit_grid",
    "nppe_system_info", "nppe_describe_system", "nppe_unknown_order", "nppe_check_csr", "nppe_solve", "nppe_solve_device", "nppe_apply",
    "nppe_memory_bytes", "nppe_profile_enable", "nppe_profile_select", "nppe_profile_read", "nppe_synth_cosine_rhs",
    "nppe_set_tolerance", "nppe_estimate_bytes", "nppe_estimate_bytes_for", "nppe_batch_solve", "nppe_server_create", "nppe_server_destroy",
    "nppe_server_add_grid", "nppe_server_submit", "nppe_server_wait", "nppe_server_last_error",
    "nppe_server_memory", "nppe_rhs_device", "nppe_rhs", "nppe_project_device", "nppe_project",
    "nppe_enforce_pressure_bc_device", "nppe_increment_reference_norm_device", "nppe_scatter_pressure_device",
    "nppe_accumulate_increment_device", "nppe_solve_laplacian_device", "nppe_solve_laplacian",
    "nppe_pressure_step_device", "nppe_projection_step_device", "nppe_update_layout", "nppe_update_layout_device",
    "nppe_increment_reference_norm_noalloc", "nppe_warm_start_guess_device", "nppe_scatter_pressure_level_device",
    "nppe_accumulate_increment_level_device", "nppe_pressure_step_ex_device", "nppe_projection_step_ex_device",
]


class NppeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"nppe error {code}: {msg}")
        self.code = code


class Options(C.Structure):
    _fields_ = [
        ("device", C.c_int32), ("method", C.c_int32), ("tolerance", C.c_double), ("max_iters", C.c_int32),
        ("compatibility", C.c_int32), ("pre_smooth", C.c_int32), ("post_smooth", C.c_int32),
        ("inner_reduction", C.c_double), ("max_refinements", C.c_int32), ("krylov", C.c_int32),
        ("hbm_budget", C.c_uint64),
    ]


class GridDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("non_uniform", C.c_int32),
        ("reserved0", C.c_int32), ("dh", C.c_double), ("cell_types", C.c_void_p), ("width_x", C.c_void_p),
        ("width_y", C.c_void_p), ("width_z", C.c_void_p),
    ]


class Result(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32), ("refinements", C.c_int32), ("status", C.c_int32), ("levels", C.c_int32),
        ("rhs_norm", C.c_double), ("residual_norm", C.c_double), ("relative_residual", C.c_double),
        ("device_ms", C.c_double), ("algorithmic_bytes", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class InflowLevel(C.Structure):
    _fields_ = [("d_level", C.c_void_p), ("centre_x", C.c_void_p), ("centre_y", C.c_void_p), ("centre_z", C.c_void_p)]


class StepOptions(C.Structure):
    _fields_ = [("d_p_prev", C.c_void_p), ("warm_start", C.c_int32), ("reserved0", C.c_int32),
                ("d_prescribed", C.c_void_p), ("inflow_level", C.c_void_p), ("warm_started", C.c_void_p)]


class SystemReport(C.Structure):
    _fields_ = [
        ("unknowns", C.c_int64), ("singular", C.c_int32), ("levels", C.c_int32), ("components", C.c_int64),
        ("floating_components", C.c_int64), ("isolated_unknowns", C.c_int64), ("coarsest_unknowns", C.c_int64),
        ("device_bytes", C.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """Loads libnppe.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `python -m nereus_cfd_amd.build` "
                                "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P, I32, I64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.nppe_default_options.argtypes = [P]
    L.nppe_default_options.restype = None
    L.nppe_create.argtypes = [P, P]
    L.nppe_create.restype = C.c_int
    L.nppe_destroy.argtypes = [P]
    L.nppe_destroy.restype = C.c_int
    L.nppe_last_error.argtypes = [P]
    L.nppe_last_error.restype = C.c_char_p
    L.nppe_commit_grid.argtypes = [P, P]
    L.nppe_commit_grid.restype = I64
    L.nppe_system_info.argtypes = [P, P, P]
    L.nppe_system_info.restype = C.c_int
    L.nppe_describe_system.argtypes = [P, P]
    L.nppe_describe_system.restype = C.c_int
    L.nppe_unknown_order.argtypes = [P, P]
    L.nppe_unknown_order.restype = C.c_int
    L.nppe_check_csr.argtypes = [P, I32, I64, P, P, P]
    L.nppe_check_csr.restype = C.c_int
    L.nppe_solve.argtypes = [P, P, P, P, D, P]
    L.nppe_solve.restype = C.c_int
    L.nppe_solve_device.argtypes = [P, P, P, P, D, P]
    L.nppe_solve_device.restype = C.c_int
    L.nppe_apply.argtypes = [P, P, P]
    L.nppe_apply.restype = C.c_int
    L.nppe_memory_bytes.argtypes = [P, P]
    L.nppe_memory_bytes.restype = C.c_int
    L.nppe_profile_enable.argtypes = [P, C.c_int]
    L.nppe_profile_enable.restype = C.c_int
    L.nppe_profile_select.argtypes = [P, C.c_uint32]
    L.nppe_profile_select.restype = C.c_int
    L.nppe_profile_read.argtypes = [P, C.c_int, P, P, P]
    L.nppe_profile_read.restype = C.c_int
    L.nppe_synth_cosine_rhs.argtypes = [P, C.c_uint64, I32, D, I32, P]
    L.nppe_synth_cosine_rhs.restype = I64
    L.nppe_set_tolerance.argtypes = [P, D]
    L.nppe_set_tolerance.restype = C.c_int
    L.nppe_estimate_bytes.argtypes = [P, I32]
    L.nppe_estimate_bytes.restype = C.c_uint64
    L.nppe_estimate_bytes_for.argtypes = [P, P]
    L.nppe_estimate_bytes_for.restype = C.c_uint64
    L.nppe_batch_solve.argtypes = [P, P, P, I32, I32, I32, P]
    L.nppe_batch_solve.restype = C.c_int
    L.nppe_server_create.argtypes = [P, P, I32, C.c_uint64]
    L.nppe_server_create.restype = C.c_int
    L.nppe_server_destroy.argtypes = [P]
    L.nppe_server_destroy.restype = C.c_int
    L.nppe_server_add_grid.argtypes = [P, C.c_uint64, P]
    L.nppe_server_add_grid.restype = C.c_int
    L.nppe_server_submit.argtypes = [P, C.c_uint64, P, D, P]
    L.nppe_server_submit.restype = I64
    L.nppe_server_wait.argtypes = [P, I64, P, P]
    L.nppe_server_wait.restype = C.c_int
    L.nppe_server_last_error.argtypes = [P]
    L.nppe_server_last_error.restype = C.c_char_p
    L.nppe_server_memory.argtypes = [P, P, P]
    L.nppe_server_memory.restype = C.c_int
    for name, args in (("nppe_rhs_device", [P, P, P, P, P, P, D, D, P]), ("nppe_rhs", [P, P, P, P, P, P, D, D, P]),
                       ("nppe_project_device", [P, P, P, P, P, D, D]), ("nppe_project", [P, P, P, P, P, D, D]),
                       ("nppe_enforce_pressure_bc_device", [P, P, P]),
                       ("nppe_increment_reference_norm_device", [P, P, P, P]),
                       ("nppe_increment_reference_norm_noalloc", [P, P, P, P]),
                       ("nppe_scatter_pressure_device", [P, P, P, P, P]),
                       ("nppe_accumulate_increment_device", [P, P, P, P]),
                       ("nppe_solve_laplacian_device", [P, P, P, P]), ("nppe_solve_laplacian", [P, P, P, P]),
                       ("nppe_pressure_step_device", [P, P, P, P, P, P, D, D, P]),
                       ("nppe_projection_step_device", [P, P, P, P, P, P, D, D, P]),
                       ("nppe_update_layout", [P, P]), ("nppe_update_layout_device", [P, P]),
                       ("nppe_warm_start_guess_device", [P, P, P, P]),
                       ("nppe_scatter_pressure_level_device", [P, P, P, P, P, P]),
                       ("nppe_accumulate_increment_level_device", [P, P, P, P, P]),
                       ("nppe_pressure_step_ex_device", [P, P, P, P, P, P, D, D, P, P]),
                       ("nppe_projection_step_ex_device", [P, P, P, P, P, P, D, D, P, P])):
        getattr(L, name).argtypes = args
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_options(**kw) -> Options:
    o = Options()
    lib().nppe_default_options(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


@dataclass
class Grid:
    """A MAC grid as CommitPressureMatrix receives it: cell types (ny, nx, nz) in MAC order and the width tables
    (ring cells included)."""
    dim: int
    types: np.ndarray
    wx: np.ndarray
    wy: np.ndarray
    wz: np.ndarray
    dh: float
    non_uniform: bool = False

    def desc(self) -> GridDesc:
        ny, nx, nz = self.types.shape
        self._keep = (np.ascontiguousarray(self.types, np.uint8), np.ascontiguousarray(self.wx, np.float64),
                      np.ascontiguousarray(self.wy, np.float64), np.ascontiguousarray(self.wz, np.float64))
        t, wx, wy, wz = self._keep
        return GridDesc(self.dim, nx, ny, nz, int(self.non_uniform), 0, float(self.dh), t.ctypes.data, wx.ctypes.data,
                        wy.ctypes.data, wz.ctypes.data)

    @property
    def unknowns(self) -> int:
        t = self.types
        inner = t[1:-1, 1:-1, 1:-1] if self.dim == 3 else t[1:-1, 1:-1, :]
        return int((inner == FLUID).sum())


def uniform_grid(n, dim=3, length=1.0, walls=SOLID) -> Grid:
    """n interior cells per axis (int or tuple (nx, ny[, nz])), spacing length/n_x, every ring cell `walls`."""
    if isinstance(n, int):
        n = (n,) * dim
    nx, ny = n[0], n[1]
    nz = n[2] if dim == 3 else 1
    dh = length / nx
    shape = (ny + 2, nx + 2, nz + 2) if dim == 3 else (ny + 2, nx + 2, 1)
    t = np.zeros(shape, np.uint8)
    t[0] = walls
    t[-1] = walls
    t[:, 0] = walls
    t[:, -1] = walls
    if dim == 3:
        t[:, :, 0] = walls
        t[:, :, -1] = walls
    return Grid(dim, t, np.full(nx + 2, dh), np.full(ny + 2, dh), np.full(nz + 2 if dim == 3 else 1, dh), dh, False)


def synth_rhs(grid: Grid, seed: int = 20261203, modes: int = 8, noise: float = 0.01, zero_mean: bool = True):
    d = grid.desc()
    b = np.zeros(max(1, grid.unknowns), np.float64)
    n = lib().nppe_synth_cosine_rhs(C.byref(d), seed, modes, noise, int(zero_mean), _ptr(b))
    if n < 0:
        raise NppeError(int(n), "synth_cosine_rhs")
    return b[:n].copy()


class Solver:
    """One backend handle (nppe_solver): create, commit a grid, solve."""

    def __init__(self, **opts):
        self.opts = default_options(**opts)
        h = C.c_void_p()
        rc = lib().nppe_create(C.byref(h), C.byref(self.opts))
        if rc != OK:
            raise NppeError(rc, lib().nppe_last_error(None).decode())
        self.h = h
        self.n = 0
        self.singular = False

    def close(self):
        if getattr(self, "h", None):
            lib().nppe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int, what: str):
        if rc < 0:
            raise NppeError(rc, f"{what}: {lib().nppe_last_error(self.h).decode()}")
        return rc

    def commit(self, grid: Grid) -> int:
        d = grid.desc()
        n = lib().nppe_commit_grid(self.h, C.byref(d))
        self._check(int(n) if n < 0 else 0, "commit_grid")
        n64, sing = C.c_int64(), C.c_int32()
        lib().nppe_system_info(self.h, C.byref(n64), C.byref(sing))
        self.n, self.singular = int(n64.value), bool(sing.value)
        self.grid = grid
        return self.n

    def update_layout(self, types) -> int:
        """nppe_update_layout: 0 unchanged, 1 updated in place, 2 committed again."""
        t = np.ascontiguousarray(types, np.uint8)
        if t.shape != self.grid.types.shape:
            raise ValueError(f"layout has shape {t.shape}, the committed box {self.grid.types.shape}")
        rc = self._check(lib().nppe_update_layout(self.h, t.ctypes.data_as(C.c_void_p)), "update_layout")
        self._after_layout(t)
        return rc

    def update_layout_device(self, types_ptr: int, types_host=None) -> int:
        rc = self._check(lib().nppe_update_layout_device(self.h, C.c_void_p(types_ptr)), "update_layout_device")
        if types_host is not None:
            self._after_layout(np.ascontiguousarray(types_host, np.uint8))
        return rc

    def _after_layout(self, t):
        n64, sing = C.c_int64(), C.c_int32()
        lib().nppe_system_info(self.h, C.byref(n64), C.byref(sing))
        self.n, self.singular = int(n64.value), bool(sing.value)
        self.grid = Grid(self.grid.dim, t.copy(), self.grid.wx, self.grid.wy, self.grid.wz, self.grid.dh,
                         self.grid.non_uniform)

    def describe(self) -> dict:
        """nppe_describe_system: components, floating (Dirichlet-free) components, isolated unknowns."""
        r = SystemReport()
        self._check(lib().nppe_describe_system(self.h, C.byref(r)), "describe_system")
        return r.as_dict()

    def order(self) -> np.ndarray:
        o = np.zeros(max(1, self.n), np.int32)
        self._check(lib().nppe_unknown_order(self.h, _ptr(o)), "unknown_order")
        return o[: self.n]

    def solve(self, b, guess=None, reference_norm=0.0):
        b = np.ascontiguousarray(b, np.float64)
        if b.shape != (self.n,):
            raise ValueError(f"b has shape {b.shape}, the system has {self.n} unknowns")
        g = None if guess is None else np.ascontiguousarray(guess, np.float64)
        x = np.zeros(max(1, self.n), np.float64)
        res = Result()
        rc = self._check(lib().nppe_solve(self.h, _ptr(b), _ptr(x), _ptr(g), float(reference_norm), C.byref(res)),
                         "solve")
        return x[: self.n], res, rc

    def solve_device(self, b_ptr: int, x_ptr: int, guess_ptr: int = 0, reference_norm=0.0):
        res = Result()
        rc = self._check(lib().nppe_solve_device(self.h, C.c_void_p(b_ptr), C.c_void_p(x_ptr),
                                                 C.c_void_p(guess_ptr) if guess_ptr else None,
                                                 float(reference_norm), C.byref(res)), "solve_device")
        return res, rc

    # ---- the pressure step around the solve (MAC fields; include/nppe.h) ----------------------------------------
    def field_sizes(self):
        """Lengths of the MAC arrays u, v, w, p of the committed grid (w is 0 in 2D)."""
        Ny, Nx, Nz = self.grid.types.shape
        if self.grid.dim == 2:
            return Ny * (Nx + 1), (Ny + 1) * Nx, 0, Ny * Nx
        return Ny * (Nx + 1) * Nz, (Ny + 1) * Nx * Nz, Ny * Nx * (Nz + 1), Ny * Nx * Nz

    def rhs(self, u, v, w, p, rho, dt, p_old=None):
        """b (IDP order) from host MAC fields: rho/dt div(u) V - scale DirichletSum, made compatible if singular."""
        a = [None if q is None else np.ascontiguousarray(q, np.float64) for q in (u, v, w, p, p_old)]
        b = np.zeros(max(1, self.n), np.float64)
        self._check(lib().nppe_rhs(self.h, *[_ptr(q) for q in a], float(rho), float(dt), _ptr(b)), "rhs")
        return b[: self.n]

    def project(self, u, v, w, p, rho, dt):
        """(u, v, w) after ProjectPressure on host MAC fields (copies)."""
        u, v = np.array(u, np.float64), np.array(v, np.float64)
        w = None if w is None else np.array(w, np.float64)
        p = np.ascontiguousarray(p, np.float64)
        self._check(lib().nppe_project(self.h, _ptr(u), _ptr(v), _ptr(w), _ptr(p), float(rho), float(dt)),
                    "project")
        return u, v, w

    def solve_laplacian(self, f):
        f = np.ascontiguousarray(f, np.float64)
        x = np.zeros(max(1, self.n), np.float64)
        res = Result()
        rc = self._check(lib().nppe_solve_laplacian(self.h, _ptr(f), _ptr(x), C.byref(res)), "solve_laplacian")
        return x[: self.n], res, rc

    # device variants take raw device addresses (0 for NULL)
    @staticmethod
    def _d(ptr):
        return C.c_void_p(ptr) if ptr else None

    def rhs_device(self, u, v, w, p, p_old, rho, dt, b):
        d = self._d
        self._check(lib().nppe_rhs_device(self.h, d(u), d(v), d(w), d(p), d(p_old), float(rho), float(dt), d(b)),
                    "rhs_device")

    def project_device(self, u, v, w, p, rho, dt):
        d = self._d
        self._check(lib().nppe_project_device(self.h, d(u), d(v), d(w), d(p), float(rho), float(dt)),
                    "project_device")

    def enforce_pressure_bc_device(self, p, prescribed=0):
        self._check(lib().nppe_enforce_pressure_bc_device(self.h, self._d(p), self._d(prescribed)),
                    "enforce_pressure_bc_device")

    def increment_reference_norm_device(self, b, p_old) -> float:
        out = C.c_double()
        self._check(lib().nppe_increment_reference_norm_device(self.h, self._d(b), self._d(p_old), C.byref(out)),
                    "increment_reference_norm_device")
        return out.value

    def increment_reference_norm_noalloc(self, b, p_old) -> float:
        """nppe_increment_reference_norm_noalloc: the same norm, no allocation (the library's fixed scratch)."""
        out = C.c_double()
        self._check(lib().nppe_increment_reference_norm_noalloc(self.h, self._d(b), self._d(p_old), C.byref(out)),
                    "increment_reference_norm_noalloc")
        return out.value

    def _level(self, level):
        """An nppe_inflow_level from (d_level pointer, cx, cy, cz host arrays), kept alive on the handle."""
        if level is None:
            return None
        ptr, cx, cy, cz = level
        arrs = [np.ascontiguousarray(a, np.float64) for a in (cx, cy, cz if cz is not None else np.zeros(1))]
        lv = InflowLevel(C.c_void_p(ptr), arrs[0].ctypes.data, arrs[1].ctypes.data, arrs[2].ctypes.data)
        self._level_keep = (arrs, lv)
        return C.byref(lv)

    def warm_start_guess_device(self, p_prev, guess) -> bool:
        built = C.c_int32()
        self._check(lib().nppe_warm_start_guess_device(self.h, self._d(p_prev), self._d(guess), C.byref(built)),
                    "warm_start_guess_device")
        return bool(built.value)

    def scatter_pressure_level_device(self, x, p, p_old=0, prescribed=0, level=None):
        self._check(lib().nppe_scatter_pressure_level_device(self.h, self._d(x), self._d(p), self._d(p_old),
                                                             self._d(prescribed), self._level(level)),
                    "scatter_pressure_level_device")

    def accumulate_increment_level_device(self, p_old, p, prescribed=0, level=None):
        self._check(lib().nppe_accumulate_increment_level_device(self.h, self._d(p_old), self._d(p),
                                                                 self._d(prescribed), self._level(level)),
                    "accumulate_increment_level_device")

    def pressure_step_ex_device(self, u, v, w, p, p_old, rho, dt, p_prev=0, warm_start=False, prescribed=0,
                                level=None):
        """nppe_pressure_step_ex_device: (result, rc, warm_started)."""
        res = Result()
        ws = C.c_int32()
        lvp = self._level(level)
        o = StepOptions(C.c_void_p(p_prev) if p_prev else None, int(bool(warm_start)), 0,
                        C.c_void_p(prescribed) if prescribed else None,
                        C.cast(lvp, C.c_void_p) if lvp is not None else None, C.cast(C.byref(ws), C.c_void_p))
        rc = self._check(lib().nppe_pressure_step_ex_device(self.h, self._d(u), self._d(v), self._d(w), self._d(p),
                                                            self._d(p_old), float(rho), float(dt), C.byref(o),
                                                            C.byref(res)), "pressure_step_ex_device")
        return res, rc, bool(ws.value)

    def projection_step_ex_device(self, u, v, w, p, p_old, rho, dt, prescribed=0, level=None):
        self._check(lib().nppe_projection_step_ex_device(self.h, self._d(u), self._d(v), self._d(w), self._d(p),
                                                         self._d(p_old), float(rho), float(dt), self._d(prescribed),
                                                         self._level(level)), "projection_step_ex_device")

    def scatter_pressure_device(self, x, p, p_old=0, prescribed=0):
        d = self._d
        self._check(lib().nppe_scatter_pressure_device(self.h, d(x), d(p), d(p_old), d(prescribed)),
                    "scatter_pressure_device")

    def accumulate_increment_device(self, p_old, p, prescribed=0):
        d = self._d
        self._check(lib().nppe_accumulate_increment_device(self.h, d(p_old), d(p), d(prescribed)),
                    "accumulate_increment_device")

    def solve_laplacian_device(self, f, x):
        res = Result()
        rc = self._check(lib().nppe_solve_laplacian_device(self.h, self._d(f), self._d(x), C.byref(res)),
                         "solve_laplacian_device")
        return res, rc

    def pressure_step_device(self, u, v, w, p, p_old, rho, dt):
        d = self._d
        res = Result()
        rc = self._check(lib().nppe_pressure_step_device(self.h, d(u), d(v), d(w), d(p), d(p_old), float(rho),
                                                         float(dt), C.byref(res)), "pressure_step_device")
        return res, rc

    def projection_step_device(self, u, v, w, p, p_old, rho, dt, prescribed=0):
        d = self._d
        self._check(lib().nppe_projection_step_device(self.h, d(u), d(v), d(w), d(p), d(p_old), float(rho),
                                                      float(dt), d(prescribed)), "projection_step_device")

    def apply(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(max(1, self.n), np.float64)
        self._check(lib().nppe_apply(self.h, _ptr(x), _ptr(y)), "apply")
        return y[: self.n]

    def check_csr(self, rowptr, col, val) -> int:
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        rc = lib().nppe_check_csr(self.h, self.n, len(col), _ptr(rowptr), _ptr(col), _ptr(val))
        return rc

    def last_error(self) -> str:
        return lib().nppe_last_error(self.h).decode()

    def memory_bytes(self) -> int:
        v = C.c_uint64()
        self._check(lib().nppe_memory_bytes(self.h, C.byref(v)), "memory_bytes")
        return int(v.value)

    def set_tolerance(self, tol: float):
        self._check(lib().nppe_set_tolerance(self.h, float(tol)), "set_tolerance")

    def profile(self, on: bool = True, kinds=None):
        """Event timing per kernel kind; `kinds` (names from KERNEL_KINDS) limits it to those kinds."""
        self._check(lib().nppe_profile_enable(self.h, int(on)), "profile_enable")
        if on and kinds is not None:
            mask = 0
            for k in kinds:
                mask |= 1 << KERNEL_KINDS.index(k)
            self._check(lib().nppe_profile_select(self.h, mask), "profile_select")

    def profile_read(self):
        out = {}
        for k, name in enumerate(KERNEL_KINDS):
            cnt, ms, by = C.c_int64(), C.c_double(), C.c_double()
            lib().nppe_profile_read(self.h, k, C.byref(cnt), C.byref(ms), C.byref(by))
            if cnt.value:
                out[name] = {"launches": int(cnt.value), "ms": ms.value, "bytes": by.value}
        return out


def estimate_bytes(grid: Grid, method: int = METHOD_MGPCG) -> int:
    d = grid.desc()
    return int(lib().nppe_estimate_bytes(C.byref(d), method))


def batch_solve(solvers, bs, workers=0):
    """Solves every (solver, b) pair concurrently (host vectors); returns [(x, Result)], rc."""
    n = len(solvers)
    hs = (C.c_void_p * n)(*[s.h for s in solvers])
    bs = [np.ascontiguousarray(b, np.float64) for b in bs]
    xs = [np.zeros(max(1, s.n), np.float64) for s in solvers]
    bp = (C.c_void_p * n)(*[b.ctypes.data for b in bs])
    xp = (C.c_void_p * n)(*[x.ctypes.data for x in xs])
    res = (Result * n)()
    rc = lib().nppe_batch_solve(C.cast(hs, C.c_void_p), C.cast(bp, C.c_void_p), C.cast(xp, C.c_void_p), n, workers,
                                0, C.cast(res, C.c_void_p))
    if rc < 0:
        raise NppeError(rc, "batch_solve")
    return [(xs[i][: solvers[i].n], res[i]) for i in range(n)], rc


def batch_solve_device(solvers, b_ptrs, x_ptrs, workers=0):
    """As batch_solve with device pointers (e.g. torch data_ptr()); returns [Result], rc."""
    n = len(solvers)
    hs = (C.c_void_p * n)(*[s.h for s in solvers])
    bp = (C.c_void_p * n)(*b_ptrs)
    xp = (C.c_void_p * n)(*x_ptrs)
    res = (Result * n)()
    rc = lib().nppe_batch_solve(C.cast(hs, C.c_void_p), C.cast(bp, C.c_void_p), C.cast(xp, C.c_void_p), n, workers,
                                1, C.cast(res, C.c_void_p))
    if rc < 0:
        raise NppeError(rc, "batch_solve_device")
    return list(res), rc


class Server:
    """nppe_server: register grids, submit solve requests at any time, wait for them (continuous batching)."""

    def __init__(self, slots=4, hbm_budget=0, **opts):
        self.opts = default_options(**opts)
        h = C.c_void_p()
        rc = lib().nppe_server_create(C.byref(h), C.byref(self.opts), slots, hbm_budget)
        if rc != OK:
            raise NppeError(rc, "server_create")
        self.h = h
        self._pending = {}

    def add_grid(self, key: int, grid: Grid):
        d = grid.desc()
        rc = lib().nppe_server_add_grid(self.h, key, C.byref(d))
        if rc != OK:
            raise NppeError(rc, "server_add_grid")

    def submit(self, key: int, b, tolerance: float) -> int:
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(max(1, len(b)), np.float64)
        t = lib().nppe_server_submit(self.h, key, b.ctypes.data if len(b) else None, float(tolerance), x.ctypes.data)
        if t < 0:
            raise NppeError(int(t), "server_submit")
        self._pending[int(t)] = (x, len(b))
        return int(t)

    def wait(self, ticket: int):
        res = Result()
        lat = C.c_double()
        rc = lib().nppe_server_wait(self.h, ticket, C.byref(res), C.byref(lat))
        x, n = self._pending.pop(ticket)
        if rc < 0:
            raise NppeError(rc, "server_wait: " + lib().nppe_server_last_error(self.h).decode())
        return x[:n], res, float(lat.value), rc

    def memory(self):
        now, peak = C.c_uint64(), C.c_uint64()
        lib().nppe_server_memory(self.h, C.byref(now), C.byref(peak))
        return int(now.value), int(peak.value)

    def close(self):
        if getattr(self, "h", None):
            lib().nppe_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
