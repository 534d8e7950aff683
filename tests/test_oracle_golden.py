"""The oracle pinned before it is trusted (CPU only): against the golden fixtures written from the compiled
reference (tests/golden/make_fixtures.py), and against what the mathematics fixes -- the linear-field exactness of
the reference's suite_pressure (:39-57, bars :166-169, :230), symmetry and A 1 = 0 (suite_units.cpp:363-371),
equal-spacing stretched == uniform up to the volume scale (suite_exact3d.cpp:364-382), and a dense solve."""
import glob
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(p).startswith(("fields_", "floor_", "level_")))


def load(path):
    return dict(np.load(path, allow_pickle=False))


def grid_of(f):
    dim = int(f["dim"])
    wz = f["wz"] if dim == 3 else np.array([float(f["dh"])])
    return O.Grid(dim, f["types"], f["wx"], f["wy"], wz, float(f["dh"]), bool(f["ref_non_uniform"]))


def refs_of(f):
    out = ([], [], [])
    for a, s, e, fac, mode in f["refinements"]:
        out[int(a)].append(O.Refinement(float(s), float(e), int(fac), int(mode)))
    return out


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_spacing_tables_bit_exact(path):
    f = load(path)
    dom = f["domain"]
    gs = int(f["grid_size"])
    dim = int(f["dim"])
    # MAC::InitializeGrid: GRID_SIZE cells across the shorter of x and y; dh = that side / GRID_SIZE.
    lx, ly = dom[1] - dom[0], dom[3] - dom[2]
    dh = min(lx, ly) / gs
    assert dh == float(f["dh"]) or bool(f["ref_non_uniform"])
    refs = refs_of(f)
    nu = bool(f["non_uniform"])
    axes = [("x", dom[0], dom[1], 0), ("y", dom[2], dom[3], 1)] + ([("z", dom[4], dom[5], 2)] if dim == 3 else [])
    for name, a0, a1, a in axes:
        nbase = int((a1 - a0) / dh + 1e-9)
        w = O.build_axis(a0, a1, nbase, dh, refs[a], nu)
        assert np.array_equal(w, f["w" + name]), name
        assert np.array_equal(O.gap_table(w), f["gap" + name]), name
        # widths sum to the domain (up to the pinned far face)
        assert abs(w[1:-1].sum() - nbase * dh) < 1e-12


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_assembly_bit_exact(path):
    f = load(path)
    s = O.assemble(grid_of(f))
    assert s.n == int(f["n"])
    assert s.singular == bool(f["singular"])
    assert np.array_equal(s.order, f["order"])
    assert np.array_equal(s.rowptr, f["rowptr"])
    assert np.array_equal(s.col, f["col"])
    assert np.array_equal(s.val, f["val"]), "CSR values differ from the reference's assembly"


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_cg_matches_reference_eigen(path):
    f = load(path)
    s = O.assemble(grid_of(f))
    b = f["rhs"]
    x, it, _ = s.cg(b, float(f["cg_tol"]), 5000)
    assert abs(it - int(f["cg_iters"])) <= 1, (it, int(f["cg_iters"]))
    ref = f["cg_x"]
    assert np.linalg.norm(x - ref) <= 1e-9 * max(np.linalg.norm(ref), 1e-300)
    # AMGCL's answer agrees with Eigen's once both constants are removed on a singular system
    amg = f["amg_x"]
    if s.singular:
        amg, x = amg - amg.mean(), x - x.mean()
    assert np.linalg.norm(amg - x) <= 1e-6 * np.linalg.norm(x)


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_matrix_invariants(path):
    f = load(path)
    s = O.assemble(grid_of(f))
    A = s.dense()
    assert np.array_equal(A, A.T), "A must be symmetric (suite_units.cpp:363-371)"
    assert np.all(np.diag(A) <= 0.0)
    # rows whose neighbours are all unknowns or Neumann: A 1 = 0 exactly
    t = f["types"]
    dim = int(f["dim"])
    row_sum = A.sum(axis=1)
    Ny, Nx, Nz = t.shape
    for r, c in enumerate(s.order):
        k, j, i = c % Nz, (c // Nz) % Nx, c // (Nz * Nx)
        nb = [(i + 1, j, k), (i - 1, j, k), (i, j + 1, k), (i, j - 1, k)]
        if dim == 3:
            nb += [(i, j, k + 1), (i, j, k - 1)]
        ring = lambda p: p[0] in (0, Ny - 1) or p[1] in (0, Nx - 1) or (dim == 3 and p[2] in (0, Nz - 1))
        dirichlet = any(t[p] in (O.EMPTY, O.PRESSURE) or (t[p] == O.FLUID and ring(p)) for p in nb)
        if not dirichlet:
            assert abs(row_sum[r]) <= 1e-12 * abs(A[r, r]) + 1e-300
        else:
            assert row_sum[r] < 0.0
    assert s.singular == bool(np.all(np.abs(row_sum) <= 1e-12 * np.abs(np.diag(A)).max()))


def test_equal_spacing_stretched_is_uniform_times_volume():
    u = load(os.path.join(os.path.dirname(__file__), "golden", "u3_neumann_8.npz"))
    st = load(os.path.join(os.path.dirname(__file__), "golden", "s3_equal_neumann.npz"))
    h = float(u["dh"])
    assert np.array_equal(u["col"], st["col"])
    # uniform: L (1/h^2 stencil); stretched form: V L with V = h^3 -- exact for h a power of two
    assert np.array_equal(u["val"] * h ** 3, st["val"])


def _linear_field_case(dim, n):
    """A box whose ring is entirely PRESSURE cells holding p = 0.3 + 1.1x - 0.7y on the boundary faces
    (suite_pressure.cpp:55-57): the interior solution of A p = b is the same linear field to round-off."""
    dh = 1.0 / n
    shape = (n + 2, n + 2, n + 2) if dim == 3 else (n + 2, n + 2, 1)
    t = np.full(shape, O.PRESSURE, np.uint8)
    if dim == 3:
        t[1:-1, 1:-1, 1:-1] = O.FLUID
    else:
        t[1:-1, 1:-1, 0] = O.FLUID
    w = np.full(n + 2, dh)
    g = O.Grid(dim, t, w, w, w if dim == 3 else np.array([dh]), dh, False)
    s = O.assemble(g)
    exact = lambda x, y: 0.3 + 1.1 * x - 0.7 * y
    Ny, Nx, Nz = shape
    centre = lambda q: (q - 0.5) * dh                   # interior cell centre
    face = lambda q, N: 0.0 if q == 0 else (1.0 if q == N - 1 else centre(q))  # a ring value sits on the face
    b = np.zeros(s.n)
    p = np.zeros(s.n)
    for r, c in enumerate(s.order):
        k, j, i = c % Nz, (c // Nz) % Nx, c // (Nz * Nx)
        p[r] = exact(centre(j), centre(i))
        nb = [(i + 1, j, k), (i - 1, j, k), (i, j + 1, k), (i, j - 1, k)]
        if dim == 3:
            nb += [(i, j, k + 1), (i, j, k - 1)]
        for (a, bb, cc) in nb:
            if t[a, bb, cc] == O.PRESSURE:
                # uniform grid: face coefficient 1, ring factor FaceGap/(Width/2) = 2, matrix scale 1/dh^2
                b[r] -= (1.0 / dh ** 2) * 2.0 * exact(face(bb, Nx), face(a, Ny))
    return s, b, p


@pytest.mark.parametrize("dim,n", [(2, 40), (3, 12)])
def test_linear_field_is_reproduced(dim, n):
    s, b, p = _linear_field_case(dim, n)
    assert np.abs(s.matvec(p) - b).max() <= 1e-9 * np.abs(b).max()
    x, it, _ = s.cg(b, 1e-12, 10000)
    assert np.abs(x - p).max() < 1e-8          # the reference's 3D Eigen bar (suite_pressure.cpp:230)
    assert np.abs(p).max() > 0.1


def test_dense_solve_small():
    f = load(os.path.join(os.path.dirname(__file__), "golden", "u3_mixed_10.npz"))
    s = O.assemble(grid_of(f))
    A = s.dense()
    b = f["rhs"]
    x = np.linalg.solve(A, b)
    xc, it, _ = s.cg(b, 1e-13, 5000)
    assert np.linalg.norm(xc - x) <= 1e-9 * np.linalg.norm(x)


def test_subtract_mean_twice():
    v = np.random.default_rng(0).standard_normal(100001) + 3.0
    w = O.compatibility(v)
    assert abs(w.mean()) < 1e-15
    assert np.allclose(w, v - v.mean(), atol=1e-12)


def test_refinement_counts_match_suite_units():
    """suite_units.cpp:319-336: GRID_SIZE 20 on [0,1], x refined 2x hard on [0.4, 0.6] -> Nx = 26 incl. ring."""
    w = O.build_axis(0.0, 1.0, 20, 0.05, [O.Refinement(0.4, 0.6, 2, O.REFINE_HARD)], True)
    assert len(w) == 26
    assert abs(w[1:-1].sum() - 1.0) < 1e-14
    u = O.build_axis(0.0, 1.0, 20, 0.05, [], False)
    assert np.all(u == 0.05) and len(u) == 22
