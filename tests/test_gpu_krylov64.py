"""MGPCG with the FP64 Krylov recursion around the FP32 V-cycle (nppe_options::krylov = NPPE_KRYLOV_FP64, the
default): the true FP64 residual meets the tolerance, the solution matches the oracle's FP64 CG and the
refinement scheme's (NPPE_KRYLOV_REFINE), repeated solves are bit-identical, and graph replay equals direct
launches."""
import numpy as np
import pytest

import oracle as O
import nereus_cfd_amd as N
import nereus_cfd_amd.workloads as W
from tests.cases import make

pytestmark = pytest.mark.gpu

REFS = ((O.Refinement(0.3, 0.6, 2, O.REFINE_HARD),), (O.Refinement(0.2, 0.8, 3, O.REFINE_SYMMETRIC),),
        (O.Refinement(0.1, 0.5, 2, O.REFINE_RAMP),))


@pytest.mark.parametrize("n,kind,stretched,tol", [
    (24, "neumann", False, 1e-8),
    (20, "outlet", False, 1e-10),
    (22, "mixed", False, 1e-9),
    (18, "mixed", True, 1e-9),
    (20, "freesurface", False, 1e-8),
])
def test_krylov64_matches_oracle(n, kind, stretched, tol):
    g, og = make(3, n, kind, non_uniform=stretched, refs=REFS if stretched else ((), (), ()), seed=4)
    s = O.assemble(og)
    b = np.random.default_rng(6).standard_normal(s.n)
    if s.singular:
        b -= b.mean()
        b -= b.mean()
    xr, _, _ = s.cg(b, 1e-13, 100000)
    with N.Solver(tolerance=tol, krylov=N.KRYLOV_FP64) as h:
        assert h.commit(g) == s.n
        x, res, rc = h.solve(b)
    assert rc == 0
    assert s.relres(b, x) < tol * 1.01 and res.relative_residual < tol
    if s.singular:
        x = x - x.mean()
        xr = xr - xr.mean()
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 2e3 * tol
    with N.Solver(tolerance=tol, krylov=N.KRYLOV_REFINE) as h2:
        h2.commit(g)
        x2, _, _ = h2.solve(b)
    if s.singular:
        x2 = x2 - x2.mean()
    assert np.linalg.norm(x - x2) / np.linalg.norm(xr) < 2e3 * tol


def test_krylov64_channel_to_1e10():
    g = W.stretched_channel(48, 24, 16)
    s = O.assemble(O.Grid(g.dim, g.types, g.wx, g.wy, g.wz, g.dh, g.non_uniform))
    b = np.random.default_rng(11).standard_normal(s.n)
    with N.Solver(tolerance=1e-10, krylov=N.KRYLOV_FP64) as h:
        h.commit(g)
        x, res, rc = h.solve(b)
    with N.Solver(tolerance=1e-10, krylov=N.KRYLOV_REFINE) as h2:
        h2.commit(g)
        _, res2, _ = h2.solve(b)
    assert rc == 0 and s.relres(b, x) < 1e-10 * 1.01
    assert res.refinements == 1 and res.iterations <= res2.iterations


def test_krylov64_deterministic_and_graph_equals_direct():
    g = N.uniform_grid(40)
    b = N.synth_rhs(g)
    with N.Solver(krylov=N.KRYLOV_FP64) as h:
        h.commit(g)
        x1, r1, _ = h.solve(b)          # graph replay
        x2, r2, _ = h.solve(b)
        h.profile(True)                 # direct launches
        x3, r3, _ = h.solve(b)
    assert r1.iterations == r2.iterations == r3.iterations
    assert np.array_equal(x1, x2) and np.array_equal(x1, x3)


@pytest.mark.parametrize("n,kind,stretched,tol", [
    (64, "neumann", False, 1e-8),
    (48, "outlet", False, 1e-10),
    (56, "mixed", False, 1e-9),
    (40, "mixed", True, 1e-9),
    (60, "freesurface", False, 1e-8),
])
def test_krylov64_2d_matches_oracle(n, kind, stretched, tol):
    """2D grids run the FP64 recursion too (the generic FP64 passes, one thread per cell): one pass to the true
    FP64 residual, the oracle's solution, and the refinement scheme's."""
    g, og = make(2, n, kind, non_uniform=stretched, refs=REFS[:2] + ((),) if stretched else ((), (), ()), seed=5)
    s = O.assemble(og)
    b = np.random.default_rng(7).standard_normal(s.n)
    if s.singular:
        b -= b.mean()
        b -= b.mean()
    xr, _, _ = s.cg(b, 1e-13, 100000)
    with N.Solver(tolerance=tol, krylov=N.KRYLOV_FP64) as h:
        assert h.commit(g) == s.n
        x, res, rc = h.solve(b)
        x1, res1, _ = h.solve(b)
    assert rc == 0 and res.refinements == 1
    assert s.relres(b, x) < tol * 1.01 and res.relative_residual < tol
    assert np.array_equal(x, x1)
    if s.singular:
        x = x - x.mean()
        xr = xr - xr.mean()
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 2e3 * tol
    with N.Solver(tolerance=tol, krylov=N.KRYLOV_REFINE) as h2:
        h2.commit(g)
        x2, _, _ = h2.solve(b)
    if s.singular:
        x2 = x2 - x2.mean()
    assert np.linalg.norm(x - x2) / np.linalg.norm(xr) < 2e3 * tol


@pytest.mark.parametrize("n,kind,batched", [(40, "neumann", False), (36, "outlet", False), (36, "outlet", True)])
def test_missed_first_pass_takes_a_second(monkeypatch, n, kind, batched):
    """The first pass's end forms only the true residual's norm. A first pass aimed at 30x the target on its
    recursive residual (NPPE_K64_AIM) ends with a true residual that misses the target; the solve then forms the
    residual vectors again from the new x and takes a second pass, which meets the tolerance -- alone and in a batch,
    where the other system takes one pass."""
    g, og = make(3, n, kind, seed=7)
    s = O.assemble(og)
    b = np.random.default_rng(8).standard_normal(s.n)
    if s.singular:
        b -= b.mean()
        b -= b.mean()
    tol = 1e-9
    with N.Solver(tolerance=tol) as h:
        h.commit(g)
        x1, r1, rc1 = h.solve(b)
    assert rc1 == 0 and r1.refinements == 1
    monkeypatch.setenv("NPPE_K64_AIM", "30")
    if not batched:
        with N.Solver(tolerance=tol) as h:
            h.commit(g)
            x2, r2, rc2 = h.solve(b)
    else:
        g2, og2 = make(3, 28, "neumann", seed=9)
        s2 = O.assemble(og2)
        b2 = np.random.default_rng(10).standard_normal(s2.n)
        b2 -= b2.mean()
        b2 -= b2.mean()
        import torch
        with N.Solver(tolerance=tol) as h, N.Solver(tolerance=tol) as h2:
            h.commit(g)
            h2.commit(g2)
            tb = [torch.from_numpy(v).cuda() for v in (b, b2)]
            tx = [torch.zeros_like(t) for t in tb]
            torch.cuda.synchronize()
            (r2, r3), rc = N.batch_solve_device([h, h2], [t.data_ptr() for t in tb], [t.data_ptr() for t in tx])
            torch.cuda.synchronize()
            assert rc == 0
            x2, x3 = (t.cpu().numpy() for t in tx)
            assert s2.relres(b2, x3) < tol * 1.01
        rc2 = 0
    assert rc2 == 0 and r2.refinements == 2, (r2.refinements, r2.iterations)
    assert r2.relative_residual < tol and s.relres(b, x2) < tol * 1.01
    if s.singular:
        x1 = x1 - x1.mean()
        x2 = x2 - x2.mean()
    assert np.linalg.norm(x2 - x1) / np.linalg.norm(x1) < 1e3 * tol


@pytest.mark.parametrize("n,kind", [(64, "neumann"), (48, "outlet"), (56, "mixed"), (60, "freesurface"),
                                    ((70, 40), "pressure")])
def test_2d_unit_coefficients_equal_general_operator(monkeypatch, n, kind):
    """A 2D grid of square cells has every FP64 face coefficient Width / FaceGap exactly 1.0 and every ring factor
    exactly 2.0, so its FP64 passes take the unit-coefficient form: MGPCG and CG64 solutions bit for bit those of
    the general operator (NPPE_UNIT2D=0)."""
    g, og = make(2, n, kind, seed=25)
    s = O.assemble(og)
    b = np.random.default_rng(26).standard_normal(s.n)
    if s.singular:
        b -= b.mean()
        b -= b.mean()
    out = {}
    for unit in ("0", "1"):
        monkeypatch.setenv("NPPE_UNIT2D", unit)
        for method in (N.METHOD_MGPCG, N.METHOD_CG64):
            with N.Solver(tolerance=1e-9, method=method) as h:
                h.commit(g)
                x, res, rc = h.solve(b)
            assert rc == 0
            out[(unit, method)] = (x, res.iterations)
    for method in (N.METHOD_MGPCG, N.METHOD_CG64):
        assert out[("0", method)][1] == out[("1", method)][1]
        assert np.array_equal(out[("0", method)][0], out[("1", method)][0])
