"""Smoother tile shapes: the autotuned shape and every forced shape (rows per tile, planes per chunk) give
bit-identical solutions, because the updates of one red-black colour do not depend on each other."""
import os

import numpy as np
import pytest

import nereus_cfd_amd as N
from tests.cases import make

pytestmark = pytest.mark.gpu

SHAPES = [("6", "1"), ("6", "5"), ("14", "3"), ("14", "16"), ("4", "2"), ("8", "7"), ("16", "1")]


def _solve(g, b, env):
    saved = {k: os.environ.get(k) for k in ("NPPE_RBGS_TM", "NPPE_RBGS_P", "NPPE_AUTOTUNE", "NPPE_RBGS_MARCH",
                                            "NPPE_FUSE_RESTRICT", "NPPE_FUSE_TM", "NPPE_FUSE_P", "NPPE_FUSE_POST",
                                            "NPPE_POST2_P", "NPPE_FUSE_PRE", "NPPE_PRE2_P", "NPPE_ROW_CODES",
                                            "NPPE_FOLD_PROLONG")}
    try:
        for k in saved:
            os.environ.pop(k, None)
        os.environ.update(env)
        with N.Solver(tolerance=1e-9) as h:
            h.commit(g)
            x, res, rc = h.solve(b)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    assert rc == 0
    return x, res


@pytest.mark.parametrize("dim,n,kind,non_uniform", [
    (3, 40, "mixed", False),             # whole-row sweep (uniform, one pair per lane), Dirichlet codes, solids
    (3, 36, "outlet", True),             # whole-row sweep on a stretched grid (width tables in registers)
    (3, 72, "neumann", False),           # large enough for the timed autotuning
    (3, (24, 20, 140), "pressure", False),   # rows of 140 cells: two pairs per lane
    (3, (20, 16, 300), "pressure", False),   # rows of 300 cells: four pairs per lane
    (2, 96, "mixed", False),             # 2D sweep
])
def test_tile_shapes_bit_identical(dim, n, kind, non_uniform):
    g, _ = make(dim, n, kind, non_uniform, seed=3)
    b = np.random.default_rng(5).standard_normal(g.unknowns)
    if kind == "neumann":
        b -= b.mean()
    ref, rres = _solve(g, b, {})
    for tm, p in SHAPES:
        x, res = _solve(g, b, {"NPPE_RBGS_TM": tm, "NPPE_RBGS_P": p})
        assert res.iterations == rres.iterations
        assert np.array_equal(x, ref), (tm, p)
    x, _ = _solve(g, b, {"NPPE_AUTOTUNE": "0"})
    assert np.array_equal(x, ref)


@pytest.mark.parametrize("n,kind,non_uniform", [
    (36, "outlet", True),                # stretched: the row sweep's register tables against the march's
    ((24, 20, 140), "mixed", False),     # two pairs per lane, Dirichlet codes
    ((20, 16, 300), "pressure", False),  # four pairs per lane
])
def test_march_equals_row_sweep(n, kind, non_uniform):
    """The marching sweep (tiles of 128 cells, forced by NPPE_RBGS_MARCH) and the whole-row sweep compute every
    update from the same values in the same order, so whole solves agree bit for bit."""
    g, _ = make(3, n, kind, non_uniform, seed=4)
    b = np.random.default_rng(6).standard_normal(g.unknowns)
    ref, rres = _solve(g, b, {})
    x, res = _solve(g, b, {"NPPE_RBGS_MARCH": "1"})
    assert res.iterations == rres.iterations
    assert np.array_equal(x, ref)


@pytest.mark.parametrize("n,kind,non_uniform", [
    (40, "mixed", False),                # one pair per lane, Dirichlet codes, solids
    ((24, 20, 140), "pressure", False),  # rows of 140 cells: two pairs per lane
    (72, "neumann", False),              # large enough for the timed autotuning
    ((33, 27, 61), "outlet", False),     # odd sizes: coarse levels with a half-width last cell (no fused form)
    ((31, 45, 97), "mixed", False),      # odd plane count: a chunk ending in a lone plane
    (36, "outlet", True),                # stretched: no fused form at any level
])
def test_fused_restriction_equals_separate_passes(n, kind, non_uniform):
    """The last pre-smoothing sweep with the restriction folded in computes every swept value and every coarse
    right-hand side with the operations of the two passes it replaces, so whole solves agree bit for bit."""
    g, _ = make(3, n, kind, non_uniform, seed=7)
    b = np.random.default_rng(8).standard_normal(g.unknowns)
    if kind == "neumann":
        b -= b.mean()
    ref, rres = _solve(g, b, {"NPPE_FUSE_RESTRICT": "0"})
    x, res = _solve(g, b, {})
    assert res.iterations == rres.iterations
    assert np.array_equal(x, ref)


@pytest.mark.parametrize("tm,p", [("6", "2"), ("8", "4"), ("10", "6"), ("12", "16"), ("12", "2"), ("6", "64")])
def test_fused_restriction_tile_shapes(tm, p):
    """Every tile shape of the fused sweep (tile rows, planes per chunk) gives the same bits (rows of 120 cells: the
    whole pre-smoothing march turned off, so that every level runs the fused sweep)."""
    g, _ = make(3, (30, 26, 120), "mixed", False, seed=9)
    b = np.random.default_rng(10).standard_normal(g.unknowns)
    ref, rres = _solve(g, b, {"NPPE_FUSE_RESTRICT": "0"})
    x, res = _solve(g, b, {"NPPE_FUSE_TM": tm, "NPPE_FUSE_P": p, "NPPE_FUSE_PRE": "0"})
    assert res.iterations == rres.iterations
    assert np.array_equal(x, ref), (tm, p)


@pytest.mark.parametrize("n,kind", [
    (40, "mixed"),                       # the coarse levels of a 40^3 grid: solids, Dirichlet codes
    ((24, 20, 140), "pressure"),         # 140-cell rows at level 0; levels 1 and 2 march both sweeps
    ((33, 27, 61), "outlet"),            # odd sizes, Dirichlet codes
    ((31, 45, 97), "mixed"),             # odd plane count, solids
    ((100, 90, 24), "neumann"),          # many rows and planes: several row tiles and chunks
])
def test_post_smoothing_pair_equals_two_sweeps(n, kind):
    """Both post-smoothing sweeps of a coarse level with rows of at most 128 cells run as one march (the second
    sweep three planes behind the first, three halo rows each side of a tile); every chunk length gives the bits of
    the two separate launches."""
    g, _ = make(3, n, kind, False, seed=11)
    b = np.random.default_rng(12).standard_normal(g.unknowns)
    if kind == "neumann":
        b -= b.mean()
    ref, rres = _solve(g, b, {"NPPE_FUSE_POST": "0"})
    for p in (None, "1", "2", "5", "64"):
        x, res = _solve(g, b, {} if p is None else {"NPPE_POST2_P": p})
        assert res.iterations == rres.iterations
        assert np.array_equal(x, ref), p


@pytest.mark.parametrize("n,kind", [
    (40, "mixed"),                       # the coarse levels of a 40^3 grid: solids, Dirichlet codes
    ((24, 20, 140), "pressure"),         # 140-cell rows at level 0; the coarser levels march
    ((33, 27, 61), "outlet"),            # odd sizes: families cut by an odd interior, Dirichlet codes
    ((31, 45, 97), "mixed"),             # odd plane count, solids
    ((100, 90, 24), "neumann"),          # many rows and planes: several row tiles and chunks
])
def test_pre_smoothing_and_restriction_equal_separate_passes(n, kind):
    """Both pre-smoothing sweeps from x = 0 and the restriction of the residual run as one march on coarse levels
    with rows of at most 128 cells (b and codes read once, no x read); every chunk length gives the bits of the
    zero-guess sweep followed by the sweep with the restriction folded in."""
    g, _ = make(3, n, kind, False, seed=16)
    b = np.random.default_rng(17).standard_normal(g.unknowns)
    if kind == "neumann":
        b -= b.mean()
    ref, rres = _solve(g, b, {"NPPE_FUSE_PRE": "0"})
    for p in (None, "2", "4", "6", "64"):
        x, res = _solve(g, b, {} if p is None else {"NPPE_PRE2_P": p})
        assert res.iterations == rres.iterations
        assert np.array_equal(x, ref), p


def test_post_smoothing_pair_batched():
    """The batched forms of the two marches (systems of different shapes in shared launches) give each system the
    bits of its own solve with the separate pre- and post-smoothing launches."""
    import torch
    cases = [make(3, 40, "mixed", False, seed=13)[0], make(3, (24, 20, 140), "pressure", False, seed=14)[0],
             make(3, 28, "neumann", False, seed=15)[0]]
    bs = []
    for i, g in enumerate(cases):
        b = np.random.default_rng(20 + i).standard_normal(g.unknowns)
        bs.append(b - b.mean() if i == 2 else b)
    refs = [_solve(g, b, {"NPPE_FUSE_POST": "0", "NPPE_FUSE_PRE": "0"})[0] for g, b in zip(cases, bs)]
    solvers = [N.Solver(tolerance=1e-9) for _ in cases]
    try:
        for h, g in zip(solvers, cases):
            h.commit(g)
        tb = [torch.from_numpy(b).cuda() for b in bs]
        tx = [torch.zeros_like(t) for t in tb]
        torch.cuda.synchronize()
        _, rc = N.batch_solve_device(solvers, [t.data_ptr() for t in tb], [t.data_ptr() for t in tx])
        torch.cuda.synchronize()
        assert rc == 0
        for t, xr in zip(tx, refs):
            assert np.array_equal(t.cpu().numpy(), xr)
    finally:
        for h in solvers:
            h.close()


@pytest.mark.parametrize("n,kind,non_uniform", [
    (40, "mixed", False),                # rows with interior solids and air cells (codes read) beside uniform rows
    ((24, 20, 140), "pressure", False),  # two pairs per lane; Dirichlet codes at every row end
    ((20, 16, 300), "outlet", False),    # four pairs per lane
    ((33, 27, 61), "neumann", False),    # odd row length: the last pair half outside the row
    (36, "outlet", True),                # stretched levels
])
def test_row_code_descriptors_equal_read_codes(n, kind, non_uniform):
    """Rows whose interior cells share one code have their codes built from a per-row descriptor instead of read
    (NPPE_ROW_CODES=0 reads them everywhere): the same codes, so the same bits."""
    g, _ = make(3, n, kind, non_uniform, seed=18)
    b = np.random.default_rng(19).standard_normal(g.unknowns)
    if kind == "neumann":
        b -= b.mean()
    ref, rres = _solve(g, b, {"NPPE_ROW_CODES": "0"})
    x, res = _solve(g, b, {})
    assert res.iterations == rres.iterations
    assert np.array_equal(x, ref)


@pytest.mark.parametrize("dim,n,kind,non_uniform", [
    (2, 96, "mixed", False),             # 2D tile sweep: solids, air cells, Dirichlet ring codes
    (2, (200, 150), "outlet", False),    # 2D, several tiles per row and odd coarse sizes
    (2, 128, "neumann", False),          # 2D pure Neumann
    (2, 80, "outlet", True),             # 2D stretched
    (3, (24, 20, 140), "mixed", False),  # 3D whole-row sweep, two pairs per lane
    (3, 36, "outlet", True),             # 3D stretched whole-row sweep
])
def test_folded_prolongation_equals_separate_pass(dim, n, kind, non_uniform):
    """The first post-smoothing sweep reads x + P x_c where it reads x (the whole-row and 2D tile sweeps) instead of
    a prolongation pass writing it first (NPPE_FOLD_PROLONG=0): the same additions, so the same bits."""
    g, _ = make(dim, n, kind, non_uniform, seed=23)
    b = np.random.default_rng(24).standard_normal(g.unknowns)
    if kind == "neumann":
        b -= b.mean()
    ref, rres = _solve(g, b, {"NPPE_FOLD_PROLONG": "0"})
    x, res = _solve(g, b, {})
    assert res.iterations == rres.iterations
    assert np.array_equal(x, ref)
