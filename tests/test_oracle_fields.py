"""The oracle's right-hand side and projection pinned before they are trusted (CPU only): bit-exact against the
fields_*.npz fixtures written from the compiled reference (MAC::GetDivergencyAt, DirichletSum,
IncrementReferenceNorm and ProjectPressure; tests/golden/make_fixtures.py), and against what the mathematics fixes:
a linear velocity field has unit divergence along its own axis, an increment against itself has no Dirichlet
term, a linear pressure has unit gradient, and a solved pressure projects the velocity to zero divergence."""
import glob
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.join(os.path.dirname(__file__), "golden")
FIELDS = sorted(glob.glob(os.path.join(HERE, "fields_*.npz")))


def load(path):
    return dict(np.load(path, allow_pickle=False))


def case_of(path):
    f = load(path)
    base = load(os.path.join(HERE, str(f["case"]) + ".npz"))
    dim = int(base["dim"])
    wz = base["wz"] if dim == 3 else np.array([float(base["dh"])])
    g = O.Grid(dim, base["types"], base["wx"], base["wy"], wz, float(base["dh"]), bool(base["ref_non_uniform"]))
    return g, O.assemble(g), f


def opt(a):
    return a if len(a) else None


@pytest.mark.parametrize("path", FIELDS, ids=[os.path.basename(p)[7:-4] for p in FIELDS])
def test_rhs_bit_exact(path):
    g, s, f = case_of(path)
    b = O.rhs(g, s, f["u"], f["v"], opt(f["w"]), f["p"], float(f["rho"]), float(f["dt"]), opt(f["p_old"]))
    assert b.shape == f["rhs"].shape
    assert np.array_equal(b, f["rhs"]), np.abs(b - f["rhs"]).max()


@pytest.mark.parametrize("path", FIELDS, ids=[os.path.basename(p)[7:-4] for p in FIELDS])
def test_projection_bit_exact(path):
    g, s, f = case_of(path)
    u, v, w = O.project(g, f["u"], f["v"], opt(f["w"]), f["p"], float(f["rho"]), float(f["dt"]))
    assert np.array_equal(u, f["u_proj"])
    assert np.array_equal(v, f["v_proj"])
    if g.dim == 3:
        assert np.array_equal(w, f["w_proj"])
    # the projection leaves the ring rows of every component alone and moves some interior face
    assert not np.array_equal(u, f["u"])


@pytest.mark.parametrize("path", FIELDS, ids=[os.path.basename(p)[7:-4] for p in FIELDS])
def test_increment_reference_norm(path):
    """IncrementReferenceNorm (PressureSolver3D.cpp:262-276) is |b + A p^n| over the unknowns."""
    g, s, f = case_of(path)
    if not len(f["p_old"]):
        assert float(f["ref_norm"]) == 0.0
        return
    pn = f["p_old"][s.order]
    ref = np.linalg.norm(f["rhs"] + s.matvec(pn))
    assert abs(ref - float(f["ref_norm"])) <= 1e-13 * ref


@pytest.mark.parametrize("path", FIELDS, ids=[os.path.basename(p)[7:-4] for p in FIELDS])
def test_increment_against_itself_has_no_dirichlet_term(path):
    g, s, f = case_of(path)
    rho, dt = float(f["rho"]), float(f["dt"])
    zeros = np.zeros_like(f["p"])
    b_same = O.rhs(g, s, f["u"], f["v"], opt(f["w"]), f["p"], rho, dt, f["p"])
    b_zero = O.rhs(g, s, f["u"], f["v"], opt(f["w"]), zeros, rho, dt)
    assert np.array_equal(b_same, b_zero)
    # the Dirichlet term is linear in p: b(p) - b(0) with no velocity scales with p
    nu, nv, nw, _ = O.field_sizes(g)
    z = [np.zeros(nu), np.zeros(nv), np.zeros(nw) if g.dim == 3 else None]
    d1 = O.rhs(g, s, *z, f["p"], rho, dt)
    d3 = O.rhs(g, s, *z, 3.0 * f["p"], rho, dt)
    assert np.allclose(d3, 3.0 * d1, rtol=1e-13, atol=1e-13 * max(1.0, np.abs(d1).max()))


def _face_positions(w):
    """Positions of the faces 0..N of an axis whose cells have widths w (an affine offset does not matter)."""
    return np.concatenate([[0.0], np.cumsum(w)])


def _centre_positions(w):
    return np.cumsum(w) - 0.5 * w


def _volumes(g):
    Ny, Nx, Nz = g.types.shape
    if not g.non_uniform:
        return lambda i, j, k: 1.0
    if g.dim == 2:
        return lambda i, j, k: g.wx[j] * g.wy[i]
    return lambda i, j, k: (g.wx[j] * g.wy[i]) * g.wz[k]


@pytest.mark.parametrize("path", FIELDS, ids=[os.path.basename(p)[7:-4] for p in FIELDS])
def test_linear_velocity_has_unit_divergence(path):
    """u = x (then v = y, w = z) on the faces: each interior cell's divergence is 1, so b = rho/dt V there."""
    g, s, f = case_of(path)
    Ny, Nx, Nz = g.types.shape
    rho, dt = 1.25, 0.004
    nu, nv, nw, npp = O.field_sizes(g)
    vol = _volumes(g)
    for axis in range(g.dim):
        # uniform grids divide by dh, stretched ones by the cell width: either way the face spacing
        w = (g.wx, g.wy, g.wz)[axis] if g.non_uniform else np.full(((Nx, Ny, Nz)[axis]), g.dh)
        xf = _face_positions(w)
        if g.dim == 3:
            u = np.zeros((Ny, Nx + 1, Nz))
            v = np.zeros((Ny + 1, Nx, Nz))
            ww = np.zeros((Ny, Nx, Nz + 1))
            (u, v, ww)[axis][...] = xf.reshape([(1, -1, 1), (-1, 1, 1), (1, 1, -1)][axis])
        else:
            u = np.zeros((Ny, Nx + 1))
            v = np.zeros((Ny + 1, Nx))
            ww = None
            (u, v)[axis][...] = xf.reshape([(1, -1), (-1, 1)][axis])
        b = O.rhs(g, s, u.ravel(), v.ravel(), None if ww is None else ww.ravel(), np.zeros(npp), rho, dt)
        expect = np.array([rho / dt * vol(c // (Nz * Nx), (c // Nz) % Nx, c % Nz) for c in s.order])
        assert np.allclose(b, expect, rtol=1e-11, atol=0.0), axis


@pytest.mark.parametrize("path", FIELDS, ids=[os.path.basename(p)[7:-4] for p in FIELDS])
def test_linear_pressure_has_unit_gradient(path):
    """p = x at the cell centres: every interior x face away from the ring and from walls moves by -dt/rho."""
    g, s, f = case_of(path)
    Ny, Nx, Nz = g.types.shape
    rho, dt = 0.8, 0.002
    nu, nv, nw, npp = O.field_sizes(g)
    wx = g.wx if g.non_uniform else np.full(Nx, g.dh)
    xc = _centre_positions(wx)
    p = np.broadcast_to(xc.reshape(1, -1, 1), (Ny, Nx, Nz)).ravel().copy()
    u0 = np.zeros(nu)
    u, v, w = O.project(g, u0, np.zeros(nv), np.zeros(nw) if g.dim == 3 else None, p, rho, dt)
    u = u.reshape(Ny, Nx + 1, Nz)
    t = g.types
    checked = 0
    krange = range(1, Nz - 1) if g.dim == 3 else range(1)
    for i in range(1, Ny - 1):
        for j in range(2, Nx - 1):
            for k in krange:
                if t[i, j - 1, k] == O.SOLID or t[i, j, k] == O.SOLID:
                    assert u[i, j, k] == 0.0
                    continue
                assert abs(u[i, j, k] + dt / rho) <= 1e-11 * dt / rho, (i, j, k)
                checked += 1
    assert checked > 0
    assert np.all(v == 0.0)


def _enforce_pressure_bc(g, p):
    """The pressure the assembled operator assumes outside the unknowns (MAC::EnforcePressureBC, MAC.cpp:866-871):
    0 in EMPTY cells and in FLUID ring cells (the zero-value Dirichlet rows), the prescribed value kept in PRESSURE
    cells, and the inward neighbour copied into every other ring cell (a Neumann row)."""
    Ny, Nx, Nz = g.types.shape
    P = p.reshape(Ny, Nx, Nz)
    t = g.types
    for i in range(Ny):
        for j in range(Nx):
            for k in range(Nz):
                ring = i in (0, Ny - 1) or j in (0, Nx - 1) or (g.dim == 3 and k in (0, Nz - 1))
                if t[i, j, k] == O.EMPTY or (ring and t[i, j, k] == O.FLUID):
                    P[i, j, k] = 0.0
                    continue
                if not ring or t[i, j, k] == O.PRESSURE:
                    continue
                ii = min(max(i, 1), Ny - 2)
                jj = min(max(j, 1), Nx - 2)
                kk = min(max(k, 1), Nz - 2) if g.dim == 3 else 0
                P[i, j, k] = P[ii, jj, kk]
    return P.ravel()


@pytest.mark.parametrize("path", FIELDS, ids=[os.path.basename(p)[7:-4] for p in FIELDS])
def test_solved_pressure_projects_to_zero_divergence(path):
    """Solve A p = b for b built from (u, p_ring), project: the divergence left at every unknown cell is round-off
    against the divergence the solve removed (the pressure step's purpose, PressureSolver3D.cpp:279-298, :417)."""
    g, s, f = case_of(path)
    if s.singular:
        pytest.skip("a singular system's right-hand side is made compatible, which changes the divergence it removes")
    rho, dt = float(f["rho"]), float(f["dt"])
    w_in = opt(f["w"])
    b = O.rhs(g, s, f["u"], f["v"], w_in, f["p"], rho, dt)
    x = np.linalg.solve(s.dense(), b)
    p = f["p"].copy()
    p[s.order] = x
    p = _enforce_pressure_bc(g, p)
    u, v, w = O.project(g, f["u"], f["v"], w_in, p, rho, dt)
    nu, nv, nw, npp = O.field_sizes(g)
    # rho/dt div V of a field = rhs(field, p) - rhs(0, p)
    z = [np.zeros(nu), np.zeros(nv), np.zeros(nw) if g.dim == 3 else None]
    div_after = O.rhs(g, s, u, v, w, p, rho, dt) - O.rhs(g, s, *z, p, rho, dt)
    div_before = O.rhs(g, s, f["u"], f["v"], w_in, p, rho, dt) - O.rhs(g, s, *z, p, rho, dt)
    assert np.abs(div_before).max() > 0.0
    assert np.abs(div_after).max() <= 1e-9 * np.abs(div_before).max(), np.abs(div_after).max()


def _ring_mask(g):
    Ny, Nx, Nz = g.types.shape
    r = np.zeros(g.types.shape, bool)
    r[0] = r[-1] = True
    r[:, 0] = r[:, -1] = True
    if g.dim == 3:
        r[:, :, 0] = r[:, :, -1] = True
    return r.ravel()


def _prescribed(g, f):
    """The boundary pressure the reference's EnforcePressureBC put on its PRESSURE ring cells (p_bc there)."""
    pr = np.zeros(g.types.size)
    m = _ring_mask(g) & (g.types.ravel() == O.PRESSURE)
    pr[m] = f["p_bc"][m]
    return pr


@pytest.mark.parametrize("path", FIELDS, ids=[os.path.basename(p)[7:-4] for p in FIELDS])
def test_enforce_pressure_bc_bit_exact(path):
    g, s, f = case_of(path)
    p = O.enforce_pressure_bc(g, f["p"], _prescribed(g, f))
    assert np.array_equal(p, f["p_bc"])
    # only ring cells change; EMPTY ring cells hold 0 and every copied ring cell holds some interior value
    ring = _ring_mask(g)
    assert np.array_equal(p[~ring], f["p"][~ring])
    t = g.types.ravel()
    assert np.all(p[ring & (t == O.EMPTY)] == 0.0)
    # with no prescribed values the PRESSURE ring cells keep theirs
    keep = O.enforce_pressure_bc(g, f["p_bc"])
    assert np.array_equal(keep, f["p_bc"])


@pytest.mark.parametrize("path", FIELDS, ids=[os.path.basename(p)[7:-4] for p in FIELDS])
def test_scatter_tail_matches_reference(path):
    g, s, f = case_of(path)
    incremental = len(f["p_old"]) > 0
    level = s.singular           # the fixtures run with the compatibility condition on
    p = O.scatter_pressure(g, s, f["x_tail"], f["p_bc"], opt(f["p_old"]), level, _prescribed(g, f))
    if level and not incremental:
        # LevelPressure's mean: Eigen's vectorised sum against the oracle's sequential one
        assert np.abs(p - f["p_tail"]).max() <= 1e-14 * np.abs(f["x_tail"]).max()
        assert not np.array_equal(p[s.order], f["x_tail"])
    else:
        assert np.array_equal(p, f["p_tail"])
    # the solve's answer sits on the unknowns (up to the mean when levelled)
    x = f["x_tail"] - (f["x_tail"].mean() if level and not incremental else 0.0)
    assert np.allclose(p[s.order], x, rtol=0, atol=1e-14 * np.abs(f["x_tail"]).max())


@pytest.mark.parametrize("path", [p for p in FIELDS if len(load(p)["p_old"])],
                         ids=[os.path.basename(p)[7:-4] for p in FIELDS if len(load(p)["p_old"])])
def test_accumulate_increment_matches_reference(path):
    g, s, f = case_of(path)
    p = O.accumulate_increment(g, s, f["p_old"], f["p_tail"], s.singular, _prescribed(g, f))
    if s.singular:
        assert np.abs(p - f["p_acc"]).max() <= 1e-14 * np.abs(f["p_acc"]).max()
    else:
        assert np.array_equal(p, f["p_acc"])
    # p^{n+1} = p^n + phi on the unknowns
    expect = f["p_old"][s.order] + f["p_tail"][s.order]
    if s.singular:
        expect = expect - expect.mean()
    assert np.allclose(p[s.order], expect, rtol=0, atol=1e-13 * np.abs(expect).max())
