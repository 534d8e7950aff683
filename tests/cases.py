"""Seeded grid cases shared by the tests: cell-type layouts the reference's assembly distinguishes (pure Neumann
box, Dirichlet outlet ring, prescribed-pressure ring, FLUID ring cells, interior solids and air pockets, free-surface
masks) on uniform and stretched spacing, 2D and 3D."""
from __future__ import annotations

import numpy as np

import oracle as O
from nereus_cfd_amd import Grid, FLUID, SOLID, EMPTY, PRESSURE, INFLOW, OUTFLOW


def axis(n, length=1.0, refs=(), non_uniform=False):
    return O.build_axis(0.0, length, n, length / n, refs, non_uniform)


def make(dim, n, kind="neumann", non_uniform=False, refs=((), (), ()), seed=0):
    """n interior cells per axis before refinement; returns (nppe Grid, oracle Grid)."""
    if isinstance(n, int):
        n = (n,) * dim
    dh = 1.0 / n[0]
    wx = axis(n[0], n[0] * dh, refs[0], non_uniform)
    wy = axis(n[1], n[1] * dh, refs[1], non_uniform)
    wz = axis(n[2], n[2] * dh, refs[2], non_uniform) if dim == 3 else np.array([dh])
    Nx, Ny, Nz = len(wx), len(wy), (len(wz) if dim == 3 else 1)
    rng = np.random.default_rng(seed)
    t = np.zeros((Ny, Nx, Nz), np.uint8)
    ring = np.zeros_like(t, bool)
    ring[0] = ring[-1] = True
    ring[:, 0] = ring[:, -1] = True
    if dim == 3:
        ring[:, :, 0] = ring[:, :, -1] = True
    t[ring] = SOLID
    if kind == "outlet":           # Dirichlet p = 0 on the x-max face
        t[:, -1, :][ring[:, -1, :]] = EMPTY
    elif kind == "pressure":       # prescribed pressure on every face
        t[ring] = PRESSURE
    elif kind == "mixed":          # every ring code, interior solids and air cells
        codes = np.array([SOLID, EMPTY, PRESSURE, INFLOW, OUTFLOW, FLUID], np.uint8)
        t[ring] = codes[rng.integers(0, len(codes), int(ring.sum()))]
        inner = ~ring
        r = rng.random(t.shape)
        t[inner & (r < 0.08)] = SOLID
        t[inner & (r >= 0.08) & (r < 0.12)] = EMPTY
    elif kind == "freesurface":    # liquid block plus droplets, air elsewhere
        inner = ~ring
        t[inner] = EMPTY
        sl = tuple(slice(1, max(2, int(0.5 * (s - 2)))) for s in t.shape[: dim])
        if dim == 3:
            t[sl[0], sl[1], 1:-1] = FLUID
        else:
            t[sl[0], sl[1], 0] = FLUID
        idx = np.argwhere(inner)
        for c in idx[rng.integers(0, len(idx), 5)]:
            t[tuple(c)] = FLUID
    g = Grid(dim, t, wx, wy, wz, dh, non_uniform)
    og = O.Grid(dim, t, wx, wy, wz, dh, non_uniform)
    return g, og
