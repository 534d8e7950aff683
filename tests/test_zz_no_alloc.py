"""nppe_increment_reference_norm_noalloc through the C ABI: the allocation-free variant of
nppe_increment_reference_norm_device (scratch in the library's fixed device array) gives the same norm bit for bit,
on two calls in a row (the second on a smaller grid, so the shared scratch holds the first call's partials beyond
the second's), leaves the handle's device memory as it was, and rejects a grid beyond its scratch."""
import glob
import os

import numpy as np
import pytest

import oracle as O
from nereus_cfd_amd import nppe as N
from tests.cases import make

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(__file__), "golden")


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).to("cuda:0")


def _norms(g, b, p_old):
    import torch
    with N.Solver() as s:
        s.commit(g)
        db, dp = _dev(b), _dev(p_old)
        torch.cuda.synchronize()
        before = s.memory_bytes()
        ref = s.increment_reference_norm_device(db.data_ptr(), dp.data_ptr())
        got = s.increment_reference_norm_noalloc(db.data_ptr(), dp.data_ptr())
        after = s.memory_bytes()
    return ref, got, before, after


@pytest.mark.parametrize("big,small", [
    ((3, 96, "mixed", False), (3, 40, "outlet", True)),
    ((3, (70, 60, 140), "freesurface", False), (2, 50, "pressure", False)),
])
def test_noalloc_matches_device_entry_point_on_two_calls(gpu, big, small):
    for dim, n, kind, nu in (big, small):
        refs = ((O.Refinement(0.3, 0.6, 2, O.REFINE_HARD),), (O.Refinement(0.2, 0.8, 3, O.REFINE_SYMMETRIC),),
                ()) if nu else ((), (), ())
        g, og = make(dim, n, kind, non_uniform=nu, refs=refs, seed=11)
        print(f"[no-alloc] {dim}D {n} {kind} stretched={nu}", flush=True)
        rng = np.random.default_rng(3)
        b = rng.standard_normal(max(1, g.unknowns))
        p_old = rng.standard_normal(g.types.size)
        ref, got, before, after = _norms(g, b, p_old)
        assert got == ref, (got, ref)
        assert before == after
        # and against the oracle's ||b + A p^n|| (the reference's IncrementReferenceNorm, PressureSolver3D.cpp:262-276)
        sys_ = O.assemble(og)
        pn = p_old[sys_.order]
        want = np.linalg.norm(b[: sys_.n] + sys_.matvec(pn))
        assert abs(got - want) <= 1e-13 * want


def test_noalloc_matches_reference_fixtures(gpu):
    for path in sorted(glob.glob(os.path.join(HERE, "fields_*.npz"))):
        f = dict(np.load(path, allow_pickle=False))
        if not len(f["p_old"]):
            continue
        base = dict(np.load(os.path.join(HERE, str(f["case"]) + ".npz"), allow_pickle=False))
        dim = int(base["dim"])
        wz = base["wz"] if dim == 3 else np.array([float(base["dh"])])
        g = N.Grid(dim, base["types"], base["wx"], base["wy"], wz, float(base["dh"]), bool(base["ref_non_uniform"]))
        ref, got, _, _ = _norms(g, f["rhs"], f["p_old"])
        assert got == ref
        assert abs(got - float(f["ref_norm"])) <= 1e-13 * float(f["ref_norm"])


def test_noalloc_rejects_grids_beyond_its_scratch(gpu):
    g = N.uniform_grid((513, 513, 1))   # 263,169 interior rows of one cell
    with N.Solver() as s:
        s.commit(g)
        b = _dev(np.ones(s.n))
        p = _dev(np.zeros(g.types.size))
        with pytest.raises(N.NppeError) as e:
            s.increment_reference_norm_noalloc(b.data_ptr(), p.data_ptr())
        assert e.value.code == N.ERR_SIZE
        assert s.increment_reference_norm_device(b.data_ptr(), p.data_ptr()) > 0.0
