"""LevelPressure's inflow shift and BuildWarmStartGuess on the device, through the C ABI, against the compiled
reference (tests/golden/level_*.npz: a scene of inlet faces with p= levels loaded by the reference's own Scene,
2D/3D, uniform/stretched, interior solids, a non-singular case; PressureSolver3D.cpp:15-30, :181-244) and the
oracle; then the extended pressure step: a warm-started Chorin step lands on the same pressure as a cold one in
fewer iterations, and the levelled pressure matches the oracle's LevelPressure of the step's own solution."""
import glob
import os

import numpy as np
import pytest

import oracle as O
from nereus_cfd_amd import nppe as N

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(__file__), "golden")
LEVEL = sorted(glob.glob(os.path.join(HERE, "level_*.npz")))
IDS = [os.path.basename(p)[:-4] for p in LEVEL]


def load(path):
    return dict(np.load(path, allow_pickle=False))


def dev(a):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a, np.float64)).to("cuda:0")
    torch.cuda.synchronize()
    return t


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def case(f):
    dim = int(f["dim"])
    g = N.Grid(dim, f["types"], f["wx"], f["wy"], f["wz"], float(f["dh"]), bool(f["non_uniform"]))
    og = O.Grid(dim, f["types"], f["wx"], f["wy"], f["wz"], float(f["dh"]), bool(f["non_uniform"]))
    return g, og, O.assemble(og)


@pytest.mark.parametrize("path", LEVEL, ids=IDS)
def test_scatter_level_matches_reference(gpu, path):
    f = load(path)
    g, og, osys = case(f)
    print(f"[level] scatter {IDS[LEVEL.index(path)]}", flush=True)
    with N.Solver() as s:
        s.commit(g)
        dl, dx, dp = dev(f["inflow_level"]), dev(f["x"]), dev(np.zeros(f["types"].size))
        lv = (dl.data_ptr(), f["cx"], f["cy"], f["cz"] if g.dim == 3 else None)
        s.scatter_pressure_level_device(dx.data_ptr(), dp.data_ptr(), 0, 0, lv)
        got = host(dp)[osys.order]
    want = f["x_levelled"]
    if not osys.singular:
        assert np.array_equal(got, want)   # LevelPressure does nothing: the scatter writes x as it is
    else:
        # the mean and the level sum reduce in another order than the reference's loops
        assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("path", LEVEL, ids=IDS)
def test_accumulate_level_matches_oracle(gpu, path):
    f = load(path)
    g, og, osys = case(f)
    rng = np.random.default_rng(5)
    p_old = rng.standard_normal(f["types"].size)
    phi = rng.standard_normal(f["types"].size)
    with N.Solver() as s:
        s.commit(g)
        dl, dpo, dp = dev(f["inflow_level"]), dev(p_old), dev(phi)
        lv = (dl.data_ptr(), f["cx"], f["cy"], f["cz"] if g.dim == 3 else None)
        s.accumulate_increment_level_device(dpo.data_ptr(), dp.data_ptr(), 0, lv)
        got = host(dp)[osys.order]
    v = (p_old + phi)[osys.order]
    if osys.singular:
        want, cnt = O.level_pressure(og, osys, v, f["inflow_level"], (f["cx"], f["cy"], f["cz"]))
        assert cnt > 0
        assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max())
    else:
        assert np.array_equal(got, v)


@pytest.mark.parametrize("path", LEVEL, ids=IDS)
def test_warm_start_guess_matches_reference(gpu, path):
    f = load(path)
    g, og, osys = case(f)
    with N.Solver() as s:
        s.commit(g)
        dpp, dg = dev(f["p_prev"]), dev(np.zeros(max(1, osys.n)))
        built = s.warm_start_guess_device(dpp.data_ptr(), dg.data_ptr())
        got = host(dg)[: osys.n]
    assert built == bool(f["warm_built"])
    if built:
        want = f["warm_guess"]
        if osys.singular:
            assert np.abs(got - want).max() <= 1e-14 * max(1.0, np.abs(want).max())
        else:
            assert np.array_equal(got, want)


def _inflow_box(n, nonuniform=False):
    """A 3D box whose x faces are INFLOW with levels (the rest SOLID): singular, levelled by its inlets."""
    from tests.cases import make
    refs = ((O.Refinement(0.3, 0.6, 2, O.REFINE_HARD),), (), ()) if nonuniform else ((), (), ())
    g, og = make(3, n, "neumann", non_uniform=nonuniform, refs=refs)
    t = g.types.copy()
    t[1:-1, 0, 1:-1] = N.INFLOW
    t[1:-1, -1, 1:-1] = N.INFLOW
    g = N.Grid(3, t, g.wx, g.wy, g.wz, g.dh, nonuniform)
    og = O.Grid(3, t, g.wx, g.wy, g.wz, g.dh, nonuniform)
    cx = np.concatenate([[-0.5 * g.wx[0]], np.cumsum(g.wx)[:-1] + 0.5 * g.wx[1:] - g.wx[0]])
    cy = np.concatenate([[-0.5 * g.wy[0]], np.cumsum(g.wy)[:-1] + 0.5 * g.wy[1:] - g.wy[0]])
    cz = np.concatenate([[-0.5 * g.wz[0]], np.cumsum(g.wz)[:-1] + 0.5 * g.wz[1:] - g.wz[0]])
    lv = np.full(t.shape, np.nan)
    lv[1:-1, 0, 1:-1] = 1.0 + cy[1:-1, None]
    lv[1:-1, -1, 1:-1] = -0.5
    return g, og, lv.ravel(), (cx, cy, cz)


@pytest.mark.parametrize("nonuniform", [False, True])
def test_pressure_step_ex_warm_start_and_level(gpu, nonuniform):
    g, og, lvals, cen = _inflow_box(24, nonuniform)
    osys = O.assemble(og)
    sizes = O.field_sizes(og)
    rng = np.random.default_rng(8)
    u, v, w = (rng.standard_normal(k) for k in sizes[:3])
    rho, dt = 1.0, 1e-3
    with N.Solver(tolerance=1e-10) as s:
        s.commit(g)
        dl = dev(lvals)
        lv = (dl.data_ptr(), *cen)
        # cold step
        du, dv, dw, dp = dev(u), dev(v), dev(w), dev(np.zeros(og.types.size))
        res0, rc0, warm0 = s.pressure_step_ex_device(du.data_ptr(), dv.data_ptr(), dw.data_ptr(), dp.data_ptr(), 0,
                                                     rho, dt, level=lv)
        p_cold = host(dp)
        assert rc0 == 0 and not warm0
        # the levelled pressure: the oracle's LevelPressure of the step's own (mean-free) solution
        x = p_cold[osys.order]
        want, cnt = O.level_pressure(og, osys, x, lvals, cen)
        assert cnt > 0 and np.abs(want - x).max() <= 1e-12 * max(1.0, np.abs(x).max())
        # warm step from the cold step's pressure, slightly perturbed velocity (a next time step)
        dpp = dev(p_cold)
        du2 = dev(u * 1.001)
        dp2 = dev(np.zeros(og.types.size))
        res1, rc1, warm1 = s.pressure_step_ex_device(du2.data_ptr(), dv.data_ptr(), dw.data_ptr(), dp2.data_ptr(), 0,
                                                     rho, dt, p_prev=dpp.data_ptr(), warm_start=True, level=lv)
        assert rc1 == 0
        assert warm1 == (not nonuniform)   # WarmStartEnabled: off on stretched grids
        # and cold on the same input, for comparison
        dp3 = dev(np.zeros(og.types.size))
        res2, _, _ = s.pressure_step_ex_device(du2.data_ptr(), dv.data_ptr(), dw.data_ptr(), dp3.data_ptr(), 0, rho,
                                               dt, level=lv)
        p_warm, p_cold2 = host(dp2), host(dp3)
    if not nonuniform:
        assert res1.iterations < res2.iterations, (res1.iterations, res2.iterations)
    scale = np.abs(p_cold2).max()
    assert np.abs(p_warm - p_cold2).max() <= 1e-7 * scale
