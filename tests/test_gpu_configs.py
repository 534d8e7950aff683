"""The benchmark configurations at their stated scope, through the C ABI:
- the stretched channel, 256 x 128 x 128 cells with BuildAxis bands (ramp / hard / ramp_back, neighbour ratio at
  most 1.05) around a voxelised cylinder: the matrix-free operator equals the oracle's assembled CSR entry by entry,
  and the solve reaches a true FP64 relative residual of 1e-10;
- 32 FLIP free-surface systems of 64^3 .. 160^3 cells (dam-break block plus spheres): the batched solve returns
  every system's alone result bit for bit, each to 1e-8;
- the serving stream: 2D grids up to 2048^2 and 3D up to 256^3, Neumann / Dirichlet outlet / free surface,
  tolerances from 1e-6 to 1e-10, through the server under an HBM budget that holds two of the largest grids: every
  served result equals the alone result bit for bit and meets its tolerance, and the device memory never exceeds the
  budget.
Residuals at full size use the handle's FP64 operator (nppe_apply), bit-identical to the oracle's CSR product
(tests/test_gpu_parity.py)."""
import time

import numpy as np
import pytest

import oracle as O
import nereus_cfd_amd as N
import nereus_cfd_amd.workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def relres(h, b, x):
    return np.linalg.norm(b - h.apply(x)) / np.linalg.norm(b)


def test_channel_256x128x128_to_1e10(gpu):
    g = W.stretched_channel()
    t0 = time.perf_counter()
    s = O.assemble(O.Grid(3, g.types, g.wx, g.wy, g.wz, g.dh, True))
    say(f"[channel] unknowns {s.n}, oracle assembly {time.perf_counter() - t0:.1f}s")
    b = np.random.default_rng(21).standard_normal(s.n)
    with N.Solver(tolerance=1e-10) as h:
        assert h.commit(g) == s.n
        assert not h.singular
        assert h.check_csr(s.rowptr, s.col, s.val) == 0, h.last_error()
        x, res, rc = h.solve(b)
        say(f"[channel] iterations {res.iterations}, passes {res.refinements}, relres {res.relative_residual:.2e}")
    assert rc == 0
    assert s.relres(b, x) < 1e-10 * 1.01


def test_batch_32_free_surface_64_to_160(gpu):
    rng = np.random.default_rng(22)
    shapes = W.batch_shapes(32, rng)
    grids = [W.flip_free_surface(shape, rng) for shape in shapes]
    bs = [W.random_rhs(gr, rng, False) for gr in grids]
    say(f"[batch] 32 grids, {sum(gr.types.size for gr in grids) / 1e6:.1f}M cells")
    solvers = [N.Solver(tolerance=1e-8) for _ in grids]
    try:
        alone = []
        for h, gr, b in zip(solvers, grids, bs):
            h.commit(gr)
            x, res, rc = h.solve(b)
            assert rc == 0
            assert relres(h, b, x) < 1e-8 * 1.01
            alone.append(x)
        out, rc = N.batch_solve(solvers, bs)
        assert rc == 0
        for (x, res), xa in zip(out, alone):
            assert np.array_equal(x, xa)
        # device vectors: the inner iterations batched into shared launches
        import torch
        tb = [torch.from_numpy(b).cuda() for b in bs]
        tx = [torch.zeros_like(t) for t in tb]
        torch.cuda.synchronize()
        res, rc = N.batch_solve_device(solvers, [t.data_ptr() for t in tb], [t.data_ptr() for t in tx])
        torch.cuda.synchronize()
        assert rc == 0 and len({r.device_ms for r in res}) == 1
        say(f"[batch] batched launches: {res[0].device_ms:.1f} ms for the 32 solves")
        for t, xa in zip(tx, alone):
            assert np.array_equal(t.cpu().numpy(), xa)
    finally:
        for h in solvers:
            h.close()


# (dim, n, bc, tolerance): every size extreme, every boundary type, both tolerance ends, plus seeded draws
REQUESTS = [(2, 2048, "neumann", 1e-6), (2, 256, "outlet", 1e-10), (2, 1024, "freesurface", 1e-8),
            (3, 256, "outlet", 1e-9), (3, 64, "freesurface", 1e-10), (3, 128, "neumann", 1e-7),
            (3, 96, "freesurface", 3e-7), (2, 512, "outlet", 2e-9)]


def test_serving_mix_at_scope(gpu):
    rng = np.random.default_rng(23)
    draws = [(d, n, bc, tol) for (d, n, bc, tol, _) in W.serving_mix(6, rng, sizes3d=(64, 96, 128))]
    reqs = REQUESTS + draws
    keys, grids, bs = {}, [], []
    for (dim, n, bc, tol) in reqs:
        k = (dim, n, bc)
        if k not in keys:
            keys[k] = len(grids)
            gr = W.serving_grid(dim, n, bc, seed=len(grids))
            grids.append(gr)
            bs.append(W.random_rhs(gr, rng, bc == "neumann"))
    say(f"[serve] {len(reqs)} requests over {len(grids)} grids")
    alone = {}
    for (dim, n, bc, tol) in reqs:
        k = keys[(dim, n, bc)]
        with N.Solver(tolerance=tol) as h:
            h.commit(grids[k])
            x, res, rc = h.solve(bs[k])
            assert rc == 0
            assert relres(h, bs[k], x) < tol * 1.01
        alone[(k, tol)] = x
    est = sorted(N.estimate_bytes(gr) for gr in grids)
    budget = int((est[-1] + est[-2]) * 1.01)
    with N.Server(slots=3, hbm_budget=budget) as srv:
        for k, gr in enumerate(grids):
            srv.add_grid(k, gr)
        tickets = []
        for (dim, n, bc, tol) in reqs:
            k = keys[(dim, n, bc)]
            tickets.append((srv.submit(k, bs[k], tol), k, tol))
        for t, k, tol in tickets:
            x, res, lat, rc = srv.wait(t)
            assert rc == 0 and lat > 0.0
            assert np.array_equal(x, alone[(k, tol)]), (k, tol)
        now, peak = srv.memory()
    say(f"[serve] peak {peak / 2**30:.2f} GiB of a {budget / 2**30:.2f} GiB budget")
    assert peak <= budget
