"""SolvePressureSystem semantics through the C ABI (PressureSettings.h:411-427, :576-624; ConjugateGradient.h:46-52):
guess, the reference-norm floor, zero and non-finite right-hand sides, the iteration cap, the HBM budget, empty
systems and re-commits."""
import numpy as np
import pytest

import oracle as O
import nereus_cfd_amd as N
from nereus_cfd_amd import nppe
from tests.cases import make

pytestmark = pytest.mark.gpu


def _sys(kind="outlet", n=12, dim=3, seed=2):
    g, og = make(dim, n, kind, seed=seed)
    return g, O.assemble(og)


@pytest.mark.parametrize("method", [N.METHOD_MGPCG, N.METHOD_CG64])
def test_exact_guess_needs_no_iterations(method):
    g, s = _sys()
    b = np.random.default_rng(1).standard_normal(s.n)
    xr, _, _ = s.cg(b, 1e-14, 10000)
    with N.Solver(method=method, tolerance=1e-8) as h:
        h.commit(g)
        x, res, rc = h.solve(b, guess=xr)
    assert rc == 0 and res.iterations == 0
    assert np.array_equal(x, xr)


@pytest.mark.parametrize("method", [N.METHOD_MGPCG, N.METHOD_CG64])
def test_reference_norm_floor_returns_the_guess(method):
    # ||b|| <= tol * referenceNorm: nothing to solve, the answer is the guess (PressureSettings.h:420-425)
    g, s = _sys()
    b = 1e-12 * np.random.default_rng(2).standard_normal(s.n)
    guess = np.random.default_rng(3).standard_normal(s.n)
    with N.Solver(method=method, tolerance=1e-8) as h:
        h.commit(g)
        x, res, rc = h.solve(b, guess=guess, reference_norm=1.0)
    assert rc == 0 and res.iterations == 0
    assert np.array_equal(x, guess)


@pytest.mark.parametrize("method", [N.METHOD_MGPCG, N.METHOD_CG64])
def test_reference_norm_sets_an_absolute_floor(method):
    g, s = _sys()
    b = np.random.default_rng(4).standard_normal(s.n)
    ref = 50.0 * np.linalg.norm(b)
    with N.Solver(method=method, tolerance=1e-6) as h:
        h.commit(g)
        x, res, rc = h.solve(b, reference_norm=ref)
    assert rc == 0
    assert np.linalg.norm(b - s.matvec(x)) < 1e-6 * ref


@pytest.mark.parametrize("method", [N.METHOD_MGPCG, N.METHOD_CG64])
def test_zero_rhs_gives_zero(method):
    g, s = _sys()
    with N.Solver(method=method) as h:
        h.commit(g)
        x, res, rc = h.solve(np.zeros(s.n), guess=np.ones(s.n))
    assert rc == 0 and res.iterations == 0 and not x.any()


def test_non_finite_rhs_is_an_error():
    g, s = _sys()
    b = np.ones(s.n)
    b[5] = np.nan
    with N.Solver() as h:
        h.commit(g)
        with pytest.raises(N.NppeError) as e:
            h.solve(b)
    assert e.value.code == nppe.ERR_NONFINITE


@pytest.mark.parametrize("method", [N.METHOD_MGPCG, N.METHOD_CG64])
def test_iteration_cap_reports_not_converged(method):
    g, s = _sys(n=16)
    b = np.random.default_rng(5).standard_normal(s.n)
    with N.Solver(method=method, tolerance=1e-12, max_iters=2) as h:
        h.commit(g)
        x, res, rc = h.solve(b)
    assert rc == nppe.NOT_CONVERGED and res.status == nppe.NOT_CONVERGED
    assert res.iterations <= 2


def test_hbm_budget_refuses_commit():
    g, _ = _sys(n=24)
    with N.Solver(hbm_budget=1 << 16) as h:
        with pytest.raises(N.NppeError) as e:
            h.commit(g)
        assert e.value.code == nppe.ERR_BUDGET
        g2, s2 = _sys(n=6)
        with N.Solver(hbm_budget=1 << 30) as h2:
            assert h2.commit(g2) == s2.n


def test_empty_system_and_recommit():
    g, s = _sys()
    empty = N.Grid(3, np.full(g.types.shape, N.SOLID, np.uint8), g.wx, g.wy, g.wz, g.dh)
    with N.Solver() as h:
        assert h.commit(empty) == 0
        x, res, rc = h.solve(np.zeros(0))
        assert rc == 0 and len(x) == 0
        assert h.commit(g) == s.n
        b = np.random.default_rng(6).standard_normal(s.n)
        x, res, rc = h.solve(b)
        assert rc == 0 and s.relres(b, x) < 1e-8


def test_solve_without_grid_is_an_error():
    with N.Solver() as h:
        with pytest.raises(N.NppeError) as e:
            h.solve(np.zeros(0))
    assert e.value.code == nppe.ERR_NO_GRID
