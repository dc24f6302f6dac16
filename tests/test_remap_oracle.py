"""The oracle's remap (remapAToBKernel, FF_SRL/FF_SRL/utils.py:32-42) pinned
by the reference's own scene: in the liver retraction scene
(liverRetraction.usd, fixture tests/golden/liver_render_map.npz) each render
point sits at the simulation vertex that mapping:pointToVertex names for it,
so remapping the rest positions reproduces the render points.  A swapped map
(src indexed by the point, dst by the vertex) or an off-by-one pair does not."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def liver():
    m = np.load(os.path.join(GOLDEN, "liver_retraction.npz"))
    r = np.load(os.path.join(GOLDEN, "liver_render_map.npz"))
    return m["mesh0_verts"], r["render_points"], r["point_to_vertex"]


def test_remap_reproduces_the_render_points(orc, liver):
    verts, points, pairs = liver
    assert pairs.shape == (points.shape[0], 2)
    assert np.array_equal(np.sort(pairs[:, 0]), np.arange(points.shape[0]))      # every point once
    got = orc.remap(np.zeros_like(points), verts, pairs)
    assert np.abs(got - points).max() <= 2e-5                 # float32 world transform of a 12-unit scene


@pytest.mark.parametrize("mistake", ["swapped", "shifted"])
def test_wrong_maps_do_not(orc, liver, mistake):
    verts, points, pairs = liver
    if mistake == "swapped":
        bad = pairs[:, ::-1] % np.array([points.shape[0], verts.shape[0]])
    else:
        bad = np.stack([pairs[:, 0], (pairs[:, 1] + 1) % verts.shape[0]], axis=1)
    got = orc.remap(np.zeros_like(points), verts, bad)
    assert np.abs(got - points).max() > 0.01


def test_last_pair_wins(orc):
    src = np.arange(12, dtype=np.float32).reshape(4, 3)
    got = orc.remap(np.zeros((2, 3), np.float32), src, [[0, 1], [1, 2], [0, 3]])
    assert np.array_equal(got, np.array([src[3], src[2]]))
