"""The benchmark workloads' host-side generators (CPU only): the product-side restatement of MAC::BuildAxis
(nereus_cfd_amd.workloads.build_axis, which builds the stretched channel's widths) reproduces the reference's
spacing tables bit for bit on every golden fixture (hard, ramp, ramp_back and symmetric bands, 2D and 3D) and the
oracle's on the channel's own band; the channel has the stated size and grading."""
import glob
import os

import numpy as np
import pytest

import oracle as O
import nereus_cfd_amd.workloads as W

GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(p).startswith(("fields_", "floor_", "level_")))


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_build_axis_matches_reference_tables(path):
    f = dict(np.load(path, allow_pickle=False))
    dom = f["domain"]
    dim = int(f["dim"])
    gs = int(f["grid_size"])
    dh = min(dom[1] - dom[0], dom[3] - dom[2]) / gs
    refs = ([], [], [])
    for a, s, e, fac, mode in f["refinements"]:
        refs[int(a)].append((float(s), float(e), int(fac), int(mode)))
    axes = [("x", dom[0], dom[1], 0), ("y", dom[2], dom[3], 1)] + ([("z", dom[4], dom[5], 2)] if dim == 3 else [])
    for name, a0, a1, a in axes:
        nbase = int((a1 - a0) / dh + 1e-9)
        w = W.build_axis(a0, a1, nbase, dh, refs[a], bool(f["non_uniform"]))
        assert np.array_equal(w, f["w" + name]), (name, np.abs(w - f["w" + name]).max() if w.shape == f["w" + name].shape else w.shape)


def test_channel_axes_match_oracle_and_size():
    wx, wy, wz, dh = W.channel_axes()
    refs = [O.Refinement(*b) for b in W.CHANNEL_BAND]
    assert np.array_equal(wx, O.build_axis(0.0, 220 / 92, 220, 1 / 92, refs, True))
    assert np.array_equal(wy, O.build_axis(0.0, 1.0, 92, 1 / 92, refs, True))
    assert (len(wx) - 2, len(wy) - 2, len(wz) - 2) == (256, 128, 128)
    for w in (wx, wy):
        inner = w[1:-1]
        ratio = np.maximum(inner[1:] / inner[:-1], inner[:-1] / inner[1:]).max()
        assert ratio <= 1.05
        assert abs(inner.min() / (dh / 2) - 1.0) < 0.05   # the factor-2 band around the cylinder (whole cells)
    g = W.stretched_channel()
    assert g.types.shape == (130, 258, 130)
    # the cylinder is voxelised: SOLID interior cells, a whole disc in every z plane
    solid = (g.types[1:-1, 1:-1, 1:-1] == O.SOLID)
    assert solid.any() and (solid == solid[:, :, :1]).all()
