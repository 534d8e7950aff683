"""nereus_cfd_amd: MI355X pressure Poisson backend for NEREUS's MAC-grid projection step.

The product is libnppe.so (C ABI in include/nppe.h; C++17 host code in csrc/ over hand-written gfx950 HIP kernels).
``nppe`` is its ctypes binding.
"""
from .nppe import (  # noqa: F401
    Grid, Solver, NppeError, uniform_grid, synth_rhs, default_options, lib, LIB_PATH,
    FLUID, SOLID, EMPTY, INFLOW, PRESSURE, OUTFLOW, METHOD_MGPCG, METHOD_CG64, KRYLOV_REFINE, KRYLOV_FP64,
    Server, batch_solve, batch_solve_device, estimate_bytes,
)
