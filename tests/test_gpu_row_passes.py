"""The whole-row marches of the FP64 Krylov passes (direction + <p, A p>, update + <r, r>) and of the residual
restriction against the 64 x 4 tile kernels they replace on 3D uniform grids (NPPE_TILE_PASSES=1 forces those).
Per cell both compute the same values (the FP64 operator in the reference's order, the FP32 residual with the same
roundings, the same family sums); only the order in which the FP64 dot products are summed differs, so the two
solves take the same iterations and agree to round-off. Layouts cover full and short rows (one and two cell pairs
per lane), Dirichlet faces on the ring and inside, free-surface masks and the pure-Neumann case."""
import os

import numpy as np
import pytest

import nereus_cfd_amd as N
from tests.cases import make

pytestmark = pytest.mark.gpu


def _solve(g, b, tiles):
    saved = os.environ.get("NPPE_TILE_PASSES")
    try:
        if tiles:
            os.environ["NPPE_TILE_PASSES"] = "1"
        else:
            os.environ.pop("NPPE_TILE_PASSES", None)
        with N.Solver(tolerance=1e-9) as h:
            h.commit(g)
            x, res, rc = h.solve(b)
    finally:
        if saved is None:
            os.environ.pop("NPPE_TILE_PASSES", None)
        else:
            os.environ["NPPE_TILE_PASSES"] = saved
    assert rc == 0
    return x, res


@pytest.mark.parametrize("n,kind,seed", [
    (40, "mixed", 3),          # short rows (20 pairs per wave), every ring code, interior solids and air
    (100, "mixed", 6),         # short rows, one pair per lane
    (128, "pressure", 4),      # full rows of one pair per lane
    (130, "outlet", 7),        # two pairs per lane, the second short; EMPTY outlet ring
    (64, "freesurface", 9),    # liquid block and droplets in air
    (72, "neumann", 5),        # pure Neumann: the mean projection
])
def test_row_passes_match_tile_passes(n, kind, seed):
    print(f"[row passes] {kind} n={n} seed={seed}", flush=True)
    g, _ = make(3, n, kind, False, seed=seed)
    b = np.random.default_rng(seed).standard_normal(g.unknowns)
    if kind == "neumann":
        b -= b.mean()
    x_row, r_row = _solve(g, b, False)
    x_tile, r_tile = _solve(g, b, True)
    assert r_row.iterations == r_tile.iterations
    assert r_row.relative_residual < 1e-9 and r_tile.relative_residual < 1e-9
    scale = np.abs(x_tile).max()
    assert np.abs(x_row - x_tile).max() <= 1e-12 * scale
