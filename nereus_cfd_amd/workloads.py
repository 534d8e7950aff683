"""Seeded synthetic pressure systems for the benchmark configurations (no counterpart in the reference, whose systems
come from a running simulation). Every generator returns nppe Grids in the reference's MAC conventions: cell types
indexed (i, j, k) = (y, x, z) with the one-cell ring, and width tables with the ring cells mirroring their
neighbours (MAC.cpp:150-161).

* stretched_channel: 3D channel 256 x 128 x 128 cells (x, y, z), a cylinder of radius 0.05 (the channel height is 1)
  across z at (0.5, 0.5), voxelised from its signed distance at the cell centres as SOLID cells; the x and y widths
  built the way the reference builds a refined axis (MAC::BuildAxis + AddRing, MAC.cpp:19-161, restated in
  build_axis): GRID_SIZE 92 base cells across the height, a factor-2 band around the cylinder -- HARD over
  [0.42, 0.58], RAMP over [0.2, 0.42] and RAMP_BACK over [0.58, 0.8] -- on both axes, the x axis 220 base cells long,
  so 256 x 128 cells with a largest neighbour ratio of 1.047; z uniform; inflow (Neumann) at x-min, a p = 0 outlet
  (EMPTY ring) at x-max, SOLID walls elsewhere; the operator is the stretched (volume-weighted) one.
* flip_free_surface: a 3D box of uniform cells with SOLID walls, a liquid dam-break block filling 30-50 % of the
  volume plus 0-200 liquid spheres of radius 2-6 cells; every other interior cell is EMPTY (p = 0, no fractional
  ghost fluid, as PressureSolver3D.cpp:162-167).
* serving_mix: request descriptions for the serving stream (2D 256^2-2048^2 or 3D 64^3-256^3; Neumann, Dirichlet
  outlet or free-surface; tolerance 1e-6 .. 1e-10).
"""
from __future__ import annotations

import numpy as np

from .nppe import Grid, FLUID, SOLID, EMPTY, INFLOW


# Refinement modes of a GridRefinement interval (the reference's Definitions.h:238-244).
REFINE_HARD, REFINE_RAMP, REFINE_SYMMETRIC, REFINE_RAMP_BACK = 0, 1, 2, 3


def _factor_at(refs, x):
    """MAC::FactorAt (MAC.cpp:19-41): the largest refinement factor of the intervals holding x; graded modes ramp
    from 1 at one edge (RAMP: start, RAMP_BACK: end, SYMMETRIC: both) to the factor."""
    f = 1.0
    for (start, end, factor, mode) in refs:
        if factor <= 1 or x < start or x > end:
            continue
        local = float(factor)
        span = end - start
        if span > 0.0 and mode != REFINE_HARD:
            t = (x - start) / span
            sn = t if mode == REFINE_RAMP else ((1.0 - t) if mode == REFINE_RAMP_BACK else (1.0 - abs(2.0 * t - 1.0)))
            local = 1.0 + (factor - 1.0) * sn
        f = max(f, local)
    return f


def build_axis(a0: float, af: float, n_base: int, dh: float, refs=(), non_uniform: bool = True) -> np.ndarray:
    """The widths of one axis with its ring cells, as MAC::BuildAxis and MAC::AddRing build them (MAC.cpp:43-161):
    refs is a sequence of (start, end, factor, mode) intervals. Uniform (or no interval): n_base cells of dh. Only
    HARD intervals: whole base cells split by the factor at their centres. Graded: the HARD interval edges anchor
    spans; a span no graded interval overlaps gets round(L f / dh) equal cells, any other is marched with the local
    target size dh / factor and rescaled to fit itself. The ring cells mirror their neighbours. The same operations in
    the same order, so the same doubles (pinned against the reference's tables by the tests)."""
    import math
    refs = [(float(s), float(e), int(f), int(m)) for (s, e, f, m) in refs]
    active = [r for r in refs if r[2] > 1]
    any_graded = any(r[3] != REFINE_HARD for r in active)
    size = []
    if not non_uniform or not active or n_base <= 0:
        size = [dh] * n_base
    elif not any_graded:
        for j in range(n_base):
            c = a0 + j * dh + dh / 2.0
            f = max(1, int(_factor_at(refs, c) + 0.5))
            size.extend([dh / float(f)] * f)
    else:
        a_end = a0 + n_base * dh
        anchors = [a0, a_end]
        for (start, end, factor, mode) in active:
            if mode == REFINE_HARD:
                for e in (start, end):
                    if a0 + 1e-12 < e < a_end - 1e-12:
                        anchors.append(e)
        edges = []
        for e in sorted(anchors):
            if not edges or e - edges[-1] > 1e-12:
                edges.append(e)
        for p, q in zip(edges[:-1], edges[1:]):
            L = q - p
            graded = any(m != REFINE_HARD and s < q - 1e-12 and e > p + 1e-12 for (s, e, f, m) in active)
            fm = _factor_at(refs, 0.5 * (p + q))
            if not graded:
                n = max(1, int(math.floor(L * fm / dh + 0.5)))
                size.extend([L / n] * n)
                continue
            first = len(size)
            x = p
            while x < q - 1e-12 and len(size) < 4000000:
                h = dh / max(1.0, _factor_at(refs, x))
                if not (h > 0.0) or not math.isfinite(h):
                    break
                size.append(h)
                x += h
            if len(size) == first:
                size.append(L)
            total = 0.0
            for t in range(first, len(size)):
                total += size[t]
            for t in range(first, len(size)):
                size[t] *= L / total
    if not size:
        return np.zeros(0)
    return np.array([size[0]] + size + [size[-1]], np.float64)


def graded_widths(n: int, length: float, centre: float, ratio: float = 1.05) -> np.ndarray:
    """n interior widths summing to length, smallest at `centre` and growing geometrically (at most `ratio` between
    neighbours) away from it, capped where the cap keeps the sum; ring widths mirror the outer cells."""
    idx = np.arange(n)
    # cell nearest to the centre under a uniform guess
    ic = int(np.clip(round(centre / length * n - 0.5), 0, n - 1))
    growth = ratio ** np.abs(idx - ic)

    def total(w0, cap):
        return np.minimum(w0 * growth, cap).sum()

    h = length / n
    # choose the cap as 1.6 uniform cells and solve for the smallest width by bisection
    cap = 1.6 * h
    lo, hi = 1e-6 * h, cap
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if total(mid, cap) < length:
            lo = mid
        else:
            hi = mid
    w = np.minimum(0.5 * (lo + hi) * growth, cap)
    w *= length / w.sum()
    return np.concatenate([[w[0]], w, [w[-1]]])


def _ring(shape, dim):
    r = np.zeros(shape, bool)
    r[0] = r[-1] = True
    r[:, 0] = r[:, -1] = True
    if dim == 3:
        r[:, :, 0] = r[:, :, -1] = True
    return r


# The channel's refinement band on the x and y axes (start, end, factor, mode), around the cylinder at 0.5.
CHANNEL_BAND = ((0.2, 0.42, 2, REFINE_RAMP), (0.42, 0.58, 2, REFINE_HARD), (0.58, 0.8, 2, REFINE_RAMP_BACK))


def channel_axes(grid_size=92, nbx=220, nz=128):
    """The channel's width tables (x, y, z), ring cells included: BuildAxis on [0, nbx / grid_size] and [0, 1] with
    CHANNEL_BAND, z uniform. The defaults give 256 x 128 x 128 cells."""
    dh = 1.0 / grid_size
    wx = build_axis(0.0, nbx / grid_size, nbx, dh, CHANNEL_BAND, True)
    wy = build_axis(0.0, 1.0, grid_size, dh, CHANNEL_BAND, True)
    wz = np.full(nz + 2, dh)
    return wx, wy, wz, dh


def stretched_channel(nx=256, ny=128, nz=128, radius=0.05, ratio=1.05) -> Grid:
    """The 256 x 128 x 128 channel (the defaults; see the module docstring). Other sizes keep the legacy geometric
    grading (graded_widths) for small test grids."""
    if (nx, ny, nz) == (256, 128, 128):
        wx, wy, wz, dh = channel_axes()
        cx, cy = 0.5, 0.5
        Ly = 1.0
    else:
        Lx, Ly, Lz = nx / ny, 1.0, nz / ny
        cx, cy = 0.25 * Lx, 0.5 * Ly
        wx = graded_widths(nx, Lx, cx, ratio)
        wy = graded_widths(ny, Ly, cy, ratio)
        wz = np.full(nz + 2, Lz / nz)
    # cell centres
    xc = np.cumsum(wx) - 0.5 * wx - wx[0]
    yc = np.cumsum(wy) - 0.5 * wy - wy[0]
    t = np.zeros((ny + 2, nx + 2, nz + 2), np.uint8)
    ring = _ring(t.shape, 3)
    t[ring] = SOLID
    t[1:-1, 0, 1:-1] = INFLOW
    t[1:-1, -1, 1:-1] = EMPTY
    sdf = np.sqrt((xc[None, :] - cx) ** 2 + (yc[:, None] - cy) ** 2) - radius    # (ny+2, nx+2)
    solid = (sdf < 0.0)
    solid[0, :] = solid[-1, :] = False
    solid[:, 0] = solid[:, -1] = False
    t[:, :, 1:-1][solid] = SOLID
    return Grid(3, t, wx, wy, wz, Ly / ny if (nx, ny, nz) != (256, 128, 128) else dh, True)


def flip_free_surface(shape, rng: np.random.Generator) -> Grid:
    nx, ny, nz = shape
    t = np.full((ny + 2, nx + 2, nz + 2), EMPTY, np.uint8)
    ring = _ring(t.shape, 3)
    t[ring] = SOLID
    fill = rng.uniform(0.3, 0.5)
    # dam-break block from the corner: full depth in z, a fraction of x and y
    fx = rng.uniform(0.5, 0.9)
    fy = min(1.0, fill / fx)
    bx, by = max(1, int(fx * nx)), max(1, int(fy * ny))
    t[1:by + 1, 1:bx + 1, 1:nz + 1] = FLUID
    nspheres = int(rng.integers(0, 201))
    if nspheres:
        zz, yy, xx = None, None, None
        for _ in range(nspheres):
            r = rng.uniform(2.0, 6.0)
            c = rng.uniform([1, 1, 1], [nx, ny, nz])
            x0, x1 = int(max(1, c[0] - r)), int(min(nx, c[0] + r)) + 1
            y0, y1 = int(max(1, c[1] - r)), int(min(ny, c[1] + r)) + 1
            z0, z1 = int(max(1, c[2] - r)), int(min(nz, c[2] + r)) + 1
            if x0 >= x1 or y0 >= y1 or z0 >= z1:
                continue
            Y, X, Z = np.meshgrid(np.arange(y0, y1), np.arange(x0, x1), np.arange(z0, z1), indexing="ij")
            inside = (X + 0.5 - 1 - c[0]) ** 2 + (Y + 0.5 - 1 - c[1]) ** 2 + (Z + 0.5 - 1 - c[2]) ** 2 <= r * r
            sub = t[y0:y1, x0:x1, z0:z1]
            sub[inside] = FLUID
    dh = 1.0 / nx
    return Grid(3, t, np.full(nx + 2, dh), np.full(ny + 2, dh), np.full(nz + 2, dh), dh, False)


def batch_shapes(count: int, rng: np.random.Generator, lo=64, hi=160):
    return [tuple(int(v) for v in rng.integers(lo, hi + 1, size=3)) for _ in range(count)]


def random_rhs(grid: Grid, rng: np.random.Generator, zero_mean: bool) -> np.ndarray:
    b = rng.standard_normal(grid.unknowns)
    if zero_mean and len(b):
        b = b - b.mean()
        b = b - b.mean()
    return b


def serving_mix(count: int, rng: np.random.Generator, sizes2d=(256, 512, 1024, 2048), sizes3d=(64, 96, 128, 192, 256)):
    """Request descriptions: (dim, n, bc, tolerance, seed). bc in {'neumann', 'outlet', 'freesurface'}."""
    out = []
    for q in range(count):
        dim = 2 if rng.random() < 0.5 else 3
        n = int(rng.choice(sizes2d if dim == 2 else sizes3d))
        bc = str(rng.choice(["neumann", "outlet", "freesurface"]))
        tol = float(10.0 ** rng.uniform(-10.0, -6.0))
        out.append((dim, n, bc, tol, int(rng.integers(0, 2 ** 31))))
    return out


def serving_grid(dim: int, n: int, bc: str, seed: int) -> Grid:
    from .nppe import uniform_grid
    if bc == "freesurface" and dim == 3:
        return flip_free_surface((n, n, n), np.random.default_rng(seed))
    g = uniform_grid(n, dim=dim)
    if bc == "outlet":
        if dim == 3:
            g.types[1:-1, -1, 1:-1] = EMPTY
        else:
            g.types[1:-1, -1, 0] = EMPTY
    elif bc == "freesurface":   # 2D: liquid below a random level, air above
        rng = np.random.default_rng(seed)
        level = int(rng.uniform(0.3, 0.7) * n)
        g.types[level + 1:-1, 1:-1, 0] = EMPTY
    return g
