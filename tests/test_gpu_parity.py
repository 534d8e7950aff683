"""GPU parity through the C ABI: the matrix-free operator against the oracle's assembled CSR, and both solvers
against the oracle's FP64 CG on the same systems (the reference's suite_pressure checks, restated)."""
import numpy as np
import pytest

import oracle as O
import nereus_cfd_amd as N
from tests.cases import make

pytestmark = pytest.mark.gpu

CASES = [
    (3, 12, "neumann", False),
    (3, 10, "outlet", False),
    (3, 9, "pressure", False),
    (3, 11, "mixed", False),
    (3, 10, "mixed", True),
    (2, 24, "neumann", False),
    (2, 21, "mixed", False),
    (2, 20, "mixed", True),
    (3, 12, "freesurface", False),
]
REFS = ((O.Refinement(0.3, 0.6, 2, O.REFINE_HARD),), (O.Refinement(0.2, 0.8, 3, O.REFINE_SYMMETRIC),),
        (O.Refinement(0.1, 0.5, 2, O.REFINE_RAMP),))


def _case(dim, n, kind, stretched, seed=1):
    return make(dim, n, kind, non_uniform=stretched, refs=REFS if stretched else ((), (), ()), seed=seed)


@pytest.mark.parametrize("dim,n,kind,stretched", CASES)
def test_operator_bitwise_matches_csr(dim, n, kind, stretched):
    g, og = _case(dim, n, kind, stretched)
    sys_ = O.assemble(og)
    with N.Solver() as s:
        assert s.commit(g) == sys_.n
        assert s.singular == sys_.singular
        assert np.array_equal(s.order(), sys_.order)
        x = np.random.default_rng(3).standard_normal(sys_.n)
        y = s.apply(x)
        assert np.array_equal(y, sys_.matvec(x)), "matrix-free FP64 apply differs from the CSR product"
        assert s.check_csr(sys_.rowptr, sys_.col, sys_.val) == 0, s.last_error()


@pytest.mark.parametrize("dim,n,kind,stretched", CASES)
@pytest.mark.parametrize("method", [N.METHOD_MGPCG, N.METHOD_CG64])
def test_solve_matches_oracle(dim, n, kind, stretched, method):
    g, og = _case(dim, n, kind, stretched)
    sys_ = O.assemble(og)
    b = np.random.default_rng(5).standard_normal(sys_.n)
    if sys_.singular:
        b = O.compatibility(b)
    xr, it, _ = sys_.cg(b, 1e-13, 20000)
    with N.Solver(method=method, tolerance=1e-10, max_iters=20000) as s:
        s.commit(g)
        x, res, rc = s.solve(b)
    assert rc == 0, s.last_error()
    assert sys_.relres(b, x) < 1e-10 * 1.01
    if sys_.singular:
        xr = xr - xr.mean()
        x = x - x.mean()
    err = np.linalg.norm(x - xr) / max(np.linalg.norm(xr), 1e-300)
    assert err < 1e-7, err


@pytest.mark.parametrize("dim,n,kind,stretched", CASES + [(3, 12, "outlet", False)])
def test_cg64_iterations_match_eigen_semantics(dim, n, kind, stretched):
    """Eigen's Jacobi-CG iteration count (the oracle's restatement, pinned to the reference's Eigen within one
    iteration) on every parity layout: 2D and 3D, stretched, singular, mixed rings and free surfaces."""
    g, og = _case(dim, n, kind, stretched)
    sys_ = O.assemble(og)
    b = np.random.default_rng(7).standard_normal(sys_.n)
    if sys_.singular:
        b = O.compatibility(b)
    xr, it, _ = sys_.cg(b, 1e-9, 5000)
    with N.Solver(method=N.METHOD_CG64, tolerance=1e-9, max_iters=5000, compatibility=0) as s:
        s.commit(g)
        x, res, rc = s.solve(b)
    assert rc == 0, s.last_error()
    assert abs(res.iterations - it) <= 1, (res.iterations, it)
    if sys_.singular:
        xr, x = xr - xr.mean(), x - x.mean()
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-7


def test_repeat_solve_is_bit_identical():
    g, og = _case(3, 16, "mixed", False)
    b = np.random.default_rng(9).standard_normal(g.unknowns)
    with N.Solver() as s:
        s.commit(g)
        x1, r1, _ = s.solve(b)
        x2, r2, _ = s.solve(b)
    assert np.array_equal(x1, x2)
    assert r1.iterations == r2.iterations
