"""The oracle's LevelPressure inflow shift and BuildWarmStartGuess pinned against the compiled reference (CPU only):
tests/golden/level_*.npz hold, per case, the scene's level at every ring cell (a scene of inlet faces with p=
expressions, loaded by the reference's own Scene), MAC::CentrePos, a vector x and what the reference's
LevelPressure made of it (PressureSolver3D.cpp:15-30, :181-211; 2D :13-29, :156-178), and a previous pressure with
BuildWarmStartGuess's answer (:215-244, including its stretched-grid and zero-field gates)."""
import glob
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(__file__)
LEVEL = sorted(glob.glob(os.path.join(HERE, "golden", "level_*.npz")))


def load(path):
    return dict(np.load(path, allow_pickle=False))


def grid_sys(f):
    g = O.Grid(int(f["dim"]), f["types"], f["wx"], f["wy"], f["wz"], float(f["dh"]), bool(f["non_uniform"]))
    return g, O.assemble(g)


def test_level_fixtures_present():
    assert len(LEVEL) >= 5


@pytest.mark.parametrize("path", LEVEL, ids=[os.path.basename(p)[:-4] for p in LEVEL])
def test_level_pressure_matches_reference(path):
    f = load(path)
    g, s = grid_sys(f)
    assert np.array_equal(s.order, f["order"])
    want = f["x_levelled"]
    if not s.singular:
        # NeedsCompatibility is false: LevelPressure leaves x as it is
        assert np.array_equal(want, f["x"])
        return
    x, cnt = O.level_pressure(g, s, f["x"], f["inflow_level"], (f["cx"], f["cy"], f["cz"]))
    assert cnt > 0
    # Eigen's vectorised mean sums in another order than the oracle's: 1e-14
    assert np.abs(x - want).max() <= 1e-14 * max(1.0, np.abs(want).max())
    # without the shift the answer is wrong by more than round-off: the fixture tests the shift itself
    x0, _ = O.level_pressure(g, s, f["x"])
    assert np.abs(x0 - want).max() > 1e-6


def test_constant_level_closed_form():
    """A level c at every inlet and an x that is linear along each inlet's normal (so the linear extrapolation to
    the ring centre is exact and equals some known value) shifts x by c minus the mean extrapolated value."""
    f = load(LEVEL[0])
    g, s = grid_sys(f)
    lv = np.where(np.isfinite(f["inflow_level"]), 1.25, np.nan)
    x = np.zeros(s.n)
    x0, cnt = O.level_pressure(g, s, x, lv, (f["cx"], f["cy"], f["cz"]))
    assert cnt > 0 and np.allclose(x0, 1.25, rtol=0, atol=1e-15)


@pytest.mark.parametrize("path", LEVEL, ids=[os.path.basename(p)[:-4] for p in LEVEL])
def test_warm_start_guess_matches_reference(path):
    f = load(path)
    g, s = grid_sys(f)
    enabled = not bool(f["non_uniform"])    # WarmStartEnabled: off on stretched grids
    built, x0 = O.warm_start_guess(g, s, f["p_prev"], s.singular)
    built = built and enabled
    assert built == bool(f["warm_built"])
    if built:
        tol = 1e-14 * max(1.0, np.abs(f["warm_guess"]).max()) if s.singular else 0.0
        assert np.abs(x0 - f["warm_guess"]).max() <= tol
