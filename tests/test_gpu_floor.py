"""NPPE_METHOD_CG64 against the compiled reference's SolveEigenPressure with a reference norm (the absolute floor
of an incremental solve, PressureSettings.h:411-427): tests/golden/floor_*.npz hold the reference's iterations
and solutions for the early return, the min(0.5, tol * ref / ||b||) branch and its cap, from zero and from a guess,
on uniform, stretched, singular and 2D systems. The handle runs without the compatibility projection, so x is
SolvePressureSystem's own answer (the reference removes the mean later, in LevelPressure)."""
import glob
import os

import numpy as np
import pytest

import nereus_cfd_amd as N

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(__file__)
FLOOR = sorted(glob.glob(os.path.join(HERE, "golden", "floor_*.npz")))


def _grid(case):
    f = dict(np.load(os.path.join(HERE, "golden", case + ".npz"), allow_pickle=False))
    dim = int(f["dim"])
    wz = f["wz"] if dim == 3 else np.array([float(f["dh"])])
    return N.Grid(dim, f["types"], f["wx"], f["wy"], wz, float(f["dh"]), bool(f["ref_non_uniform"]))


@pytest.mark.parametrize("path", FLOOR, ids=[os.path.basename(p)[:-4] for p in FLOOR])
def test_cg64_floor_matches_reference(gpu, path):
    f = dict(np.load(path, allow_pickle=False))
    case = str(f["case"])
    b, tol, guess = f["rhs"], float(f["tol"]), f["guess"]
    print(f"floor case {case}", flush=True)
    with N.Solver(method=N.METHOD_CG64, tolerance=tol, max_iters=5000, compatibility=0) as s:
        assert s.commit(_grid(case)) == b.size
        for q, name in enumerate(f["names"]):
            name = str(name)
            g = guess if int(f["has_guess"][q]) else None
            x, res, rc = s.solve(b, guess=g, reference_norm=float(f["ref_norms"][q]))
            assert rc == 0, (name, s.last_error())
            want_it, want_x = int(f["iters"][q]), f["x"][q]
            if name.startswith("early"):
                assert res.iterations == 0 and want_it == 0, name
                assert np.array_equal(x, want_x), name   # the guess, or 0, returned as it is
                continue
            assert abs(res.iterations - want_it) <= 1, (name, res.iterations, want_it)
            scale = np.linalg.norm(want_x)
            assert np.linalg.norm(x - want_x) <= 1e-8 * scale, (name, np.linalg.norm(x - want_x) / scale)
