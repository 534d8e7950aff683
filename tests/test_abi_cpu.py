"""CPU checks of the C ABI library: it loads, exports every symbol the headers declare, reports errors without a
device, and the host-side helpers behave."""
import glob
import os
import re

import numpy as np
import pytest

import nereus_cfd_amd as N
from nereus_cfd_amd import nppe

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        names.update(re.findall(r"\b(nppe_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = N.lib()
    names = declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(nppe.EXPORTS) == names


def test_default_options():
    o = N.default_options()
    assert o.method == N.METHOD_MGPCG and o.tolerance == 1e-8 and o.max_iters == 2000
    assert o.compatibility == 1 and o.pre_smooth == -1 and o.post_smooth == -1 and o.max_refinements == 8
    assert o.krylov == N.KRYLOV_FP64


def test_create_reports_errors():
    import ctypes as C
    lib = N.lib()
    assert lib.nppe_create(None, None) == nppe.ERR_ARGUMENT
    assert b"null" in lib.nppe_last_error(None)
    bad = N.default_options(tolerance=-1.0)
    h = C.c_void_p()
    assert lib.nppe_create(C.byref(h), C.byref(bad)) == nppe.ERR_ARGUMENT
    assert lib.nppe_destroy(None) == 0
    assert lib.nppe_commit_grid(None, None) == nppe.ERR_ARGUMENT
    assert lib.nppe_solve(None, None, None, None, 0.0, None) == nppe.ERR_ARGUMENT


def test_synthetic_rhs_is_seeded_and_mean_free():
    g = N.uniform_grid((10, 12, 9))
    b1 = N.synth_rhs(g, seed=5)
    b2 = N.synth_rhs(g, seed=5)
    b3 = N.synth_rhs(g, seed=6)
    assert len(b1) == g.unknowns == 10 * 12 * 9
    assert np.array_equal(b1, b2) and not np.array_equal(b1, b3)
    assert abs(b1.mean()) < 1e-15
    g2 = N.uniform_grid(16, dim=2)
    assert len(N.synth_rhs(g2)) == 256


def test_grid_helper_matches_reference_ring_convention():
    g = N.uniform_grid(4)
    assert g.types.shape == (6, 6, 6)
    assert (g.types[0] == N.SOLID).all() and (g.types[1:-1, 1:-1, 1:-1] == N.FLUID).all()
    assert np.all(g.wx == 0.25)
