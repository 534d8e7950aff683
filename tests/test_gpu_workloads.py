"""The benchmark configurations' systems, small, against the oracle; and many systems at once: batched solves and
the solve server return, for every system, exactly what the same solve returns alone (the north star's
alone-vs-in-the-mix check), with the HBM budget enforced by admission and eviction."""
import numpy as np
import pytest

import oracle as O
import nereus_cfd_amd as N
import nereus_cfd_amd.workloads as W

pytestmark = pytest.mark.gpu


def _oracle(g):
    return O.assemble(O.Grid(g.dim, g.types, g.wx, g.wy, g.wz, g.dh, g.non_uniform))


def test_stretched_channel_to_1e10():
    g = W.stretched_channel(48, 24, 16)
    s = _oracle(g)
    b = np.random.default_rng(11).standard_normal(s.n)
    xr, _, _ = s.cg(b, 1e-13, 50000)
    with N.Solver(tolerance=1e-10) as h:
        assert h.commit(g) == s.n
        assert h.check_csr(s.rowptr, s.col, s.val) == 0, h.last_error()
        x, res, rc = h.solve(b)
    assert rc == 0 and s.relres(b, x) < 1e-10 * 1.01
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-7


def test_flip_free_surface_to_1e8():
    g = W.flip_free_surface((40, 36, 28), np.random.default_rng(12))
    s = _oracle(g)
    b = np.random.default_rng(13).standard_normal(s.n)
    xr, _, _ = s.cg(b, 1e-13, 50000)
    with N.Solver(tolerance=1e-8) as h:
        assert h.commit(g) == s.n
        x, res, rc = h.solve(b)
    assert rc == 0 and s.relres(b, x) < 1e-8 * 1.01
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-5


def _mix(count=6, seed=14):
    rng = np.random.default_rng(seed)
    grids = [W.flip_free_surface(shape, rng) for shape in W.batch_shapes(count, rng, 24, 48)]
    bs = [W.random_rhs(g, rng, False) for g in grids]
    return grids, bs


def test_batch_equals_alone():
    grids, bs = _mix()
    alone = []
    for g, b in zip(grids, bs):
        with N.Solver() as h:
            h.commit(g)
            x, res, rc = h.solve(b)
            assert rc == 0
            alone.append(x)
    solvers = [N.Solver() for _ in grids]
    for h, g in zip(solvers, grids):
        h.commit(g)
    out, rc = N.batch_solve(solvers, bs)
    assert rc == 0
    for (x, res), xa in zip(out, alone):
        assert np.array_equal(x, xa)
    for h in solvers:
        h.close()


def test_graph_replay_equals_direct_launches():
    g = N.uniform_grid(40)
    b = N.synth_rhs(g)
    with N.Solver() as h:
        h.commit(g)
        x1, r1, _ = h.solve(b)          # inner iterations replayed from the captured graph
        h.profile(True)                 # profiling launches every kernel directly (no graph)
        x2, r2, _ = h.solve(b)
    assert r1.iterations == r2.iterations
    assert np.array_equal(x1, x2)


def test_server_equals_alone_and_respects_budget():
    grids, bs = _mix(5, 15)
    alone = []
    for g, b in zip(grids, bs):
        with N.Solver(tolerance=1e-9) as h:
            h.commit(g)
            alone.append(h.solve(b)[0])
    est = [N.estimate_bytes(g) for g in grids]
    budget = int(sum(sorted(est)[-2:]) * 1.01)     # room for the two largest grids at once
    with N.Server(slots=3, hbm_budget=budget) as srv:
        for k, g in enumerate(grids):
            srv.add_grid(k, g)
        tickets = [srv.submit(k % len(grids), bs[k % len(grids)], 1e-9) for k in range(3 * len(grids))]
        for q, t in enumerate(tickets):
            x, res, lat, rc = srv.wait(t)
            assert rc == 0 and lat > 0.0
            assert np.array_equal(x, alone[q % len(grids)])
        now, peak = srv.memory()
        assert peak <= budget
