"""The benchmark configurations' systems, small, against the oracle; and many systems at once: batched solves and
the solve server return, for every system, exactly what the same solve returns alone (a system's result does not
depend on what it shares launches with), with the HBM budget enforced by admission and eviction."""
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


def _mix(count=6, seed=14, shapes=None):
    rng = np.random.default_rng(seed)
    drawn = W.batch_shapes(count, rng, 24, 48)
    grids = [W.flip_free_surface(shape, rng) for shape in drawn]
    bs = [W.random_rhs(g, rng, False) for g in grids]
    if shapes is not None:
        shapes.extend(drawn)
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


def _device_batch(solvers, bs):
    import torch
    tb = [torch.from_numpy(b).cuda() for b in bs]
    tx = [torch.zeros_like(t) for t in tb]
    torch.cuda.synchronize()
    res, rc = N.batch_solve_device(solvers, [t.data_ptr() for t in tb], [t.data_ptr() for t in tx])
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in tx], res, rc


def _alone(grids, bs):
    out = []
    for g, b in zip(grids, bs):
        with N.Solver() as h:
            h.commit(g)
            x, res, rc = h.solve(b)
            assert rc == 0
            out.append((x, res))
    return out


def test_batched_launches_equal_alone():
    """Device vectors: the systems' inner iterations share launches (one per V-cycle place and kernel); every result
    is the alone result bit for bit, on the call that builds the batch's plan, on the call that reuses it, and after
    one system's layout changed (the plan is rebuilt)."""
    shapes = []
    grids, bs = _mix(8, 16, shapes)
    alone = _alone(grids, bs)
    solvers = [N.Solver() for _ in grids]
    try:
        for h, g in zip(solvers, grids):
            h.commit(g)
        for call in range(2):
            xs, res, rc = _device_batch(solvers, bs)
            assert rc == 0
            # one batch on one stream: every system reports the batch's device time
            assert len({r.device_ms for r in res}) == 1
            for x, r, (xa, ra) in zip(xs, res, alone):
                assert np.array_equal(x, xa), call
                assert r.iterations == ra.iterations and r.relative_residual == ra.relative_residual
        g2 = W.flip_free_surface(shapes[3], np.random.default_rng(99))
        solvers[3].update_layout(g2.types)
        b2 = W.random_rhs(g2, np.random.default_rng(98), False)
        assert b2.size == solvers[3].n
        grids[3], bs[3] = g2, b2
        alone[3] = _alone([g2], [b2])[0]
        xs, res, rc = _device_batch(solvers, bs)
        assert rc == 0
        for x, (xa, _) in zip(xs, alone):
            assert np.array_equal(x, xa)
    finally:
        for h in solvers:
            h.close()


def test_batched_launches_fall_back_for_other_systems():
    """A batch holding a system without the batched kernels (a 2D grid) is solved system by system, still equal to
    alone."""
    grids, bs = _mix(3, 17)
    g2 = W.serving_grid(2, 96, "outlet", seed=1)
    grids.append(g2)
    bs.append(W.random_rhs(g2, np.random.default_rng(5), False))
    alone = _alone(grids, bs)
    solvers = [N.Solver() for _ in grids]
    try:
        for h, g in zip(solvers, grids):
            h.commit(g)
        xs, res, rc = _device_batch(solvers, bs)
        assert rc == 0
        for x, (xa, _) in zip(xs, alone):
            assert np.array_equal(x, xa)
    finally:
        for h in solvers:
            h.close()


@pytest.mark.parametrize("other", ["cg64", "profiling", "stretched"])
def test_batched_launches_with_a_system_of_another_kind(other):
    """A CG64 handle, a handle that is profiling, or a stretched grid in the batch: every system is solved on its
    own stream instead, and still gets its alone result."""
    grids, bs = _mix(3, 18)
    solvers = [N.Solver() for _ in grids]
    if other == "stretched":
        g = W.stretched_channel(32, 16, 16)
        grids.append(g)
        bs.append(W.random_rhs(g, np.random.default_rng(3), False))
        solvers.append(N.Solver())
    else:
        grids.append(grids[0])
        bs.append(bs[0])
        solvers.append(N.Solver(method=N.METHOD_CG64) if other == "cg64" else N.Solver())
    alone = []
    for g, b, h in zip(grids, bs, solvers):
        with N.Solver(method=h.opts.method) as a:
            a.commit(g)
            x, res, rc = a.solve(b)
            assert rc == 0
            alone.append(x)
    try:
        for h, g in zip(solvers, grids):
            h.commit(g)
        if other == "profiling":
            solvers[-1].profile(True)
        xs, res, rc = _device_batch(solvers, bs)
        assert rc == 0
        for x, xa in zip(xs, alone):
            assert np.array_equal(x, xa)
    finally:
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
