"""The oracle's SolveEigenPressure floor (oracle_solve_pressure; PressureSettings.h:411-427) pinned against the
compiled reference (CPU only): tests/golden/floor_*.npz hold SolvePressureSystem's iterations and solutions with a
reference norm on uniform, stretched, singular and 2D systems -- the early return (||b|| <= tol * ref: the guess,
or 0, with 0 iterations), the min(0.5, tol * ref / ||b||) branch and its 0.5 cap, from zero and from a guess.

The smooth right-hand sides leave more than half of ||b|| after one CG step, so the cap cases run at least one
iteration at 0.5 where an uncapped tolerance (0.99) would stop at once: a dropped or wrong cap fails here, as does
a floor taken relative to ||b|| instead of the reference norm (the min cases would stop at 1e-10 instead)."""
import glob
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(__file__)
FLOOR = sorted(glob.glob(os.path.join(HERE, "golden", "floor_*.npz")))


def load(path):
    return dict(np.load(path, allow_pickle=False))


def system_of(case):
    f = dict(np.load(os.path.join(HERE, "golden", case + ".npz"), allow_pickle=False))
    dim = int(f["dim"])
    wz = f["wz"] if dim == 3 else np.array([float(f["dh"])])
    return O.assemble(O.Grid(dim, f["types"], f["wx"], f["wy"], wz, float(f["dh"]), bool(f["ref_non_uniform"])))


def test_floor_fixtures_present():
    assert len(FLOOR) >= 5


@pytest.mark.parametrize("path", FLOOR, ids=[os.path.basename(p)[:-4] for p in FLOOR])
def test_solve_pressure_matches_reference(path):
    f = load(path)
    s = system_of(str(f["case"]))
    b, tol, guess = f["rhs"], float(f["tol"]), f["guess"]
    for q, name in enumerate(f["names"]):
        ref = float(f["ref_norms"][q])
        g = guess if int(f["has_guess"][q]) else None
        x, it = s.solve_pressure(b, tol, 5000, guess=g, reference_norm=ref)
        want_it, want_x = int(f["iters"][q]), f["x"][q]
        if str(name).startswith("early"):
            # the early return: exactly the guess (or 0), no iterations
            assert it == 0 and want_it == 0, name
            assert np.array_equal(x, g if g is not None else np.zeros_like(b)), name
            assert np.array_equal(want_x, x), name
            continue
        assert abs(it - want_it) <= 1, (name, it, want_it)
        scale = max(np.linalg.norm(want_x), 1e-300)
        # Eigen's vectorised dot products sum in another order than the oracle's: the iterates drift by round-off
        assert np.linalg.norm(x - want_x) <= 1e-8 * scale, (name, np.linalg.norm(x - want_x) / scale)


@pytest.mark.parametrize("path", FLOOR, ids=[os.path.basename(p)[:-4] for p in FLOOR])
def test_floor_branches_are_exercised(path):
    """The fixtures cover what they claim: each branch's effective relative tolerance, from the recorded norms."""
    f = load(path)
    b, tol = f["rhs"], float(f["tol"])
    bn = np.linalg.norm(b)
    for q, name in enumerate(f["names"]):
        ref = float(f["ref_norms"][q])
        name = str(name)
        if name.startswith("early"):
            assert bn <= tol * ref
        elif name.startswith("min"):
            assert tol * ref / bn < 0.5 and bn > tol * ref
        elif name.startswith("cap"):
            assert 0.5 < tol * ref / bn <= 1.0
            assert int(f["iters"][q]) >= 1, "a cap case must run past the first step to tell 0.5 from 0.99"
        else:
            assert ref == 0.0


def test_planted_cap_mistake_is_caught():
    """A floor mapping without the 0.5 cap (tolerance tol * ref / ||b|| as it is) stops earlier on a cap case."""
    f = load(FLOOR[0])
    s = system_of(str(f["case"]))
    b, tol = f["rhs"], float(f["tol"])
    q = list(map(str, f["names"])).index("cap_zero")
    ref = float(f["ref_norms"][q])
    uncapped = tol * ref / np.linalg.norm(b)
    _, it_wrong = s.solve_pressure(b, uncapped, 5000)   # reference_norm 0: plain relative tolerance
    assert it_wrong != int(f["iters"][q])
