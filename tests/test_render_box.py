"""The renderer's fast ray / box decision (ff-srl_amd/csrc/render_box.hpp:
visit_node, slab values from reciprocal products with an exact fallback near
ties) against the exact form, the reference's intersectAABB
(render.py:588-601) with the traversal's "tmin < tMin" (render.py:447-452),
on the host: tests/render_box_harness.cpp compiled with the library's own
flags, 6 million seeded random and adversarial cases (zero and tiny
direction components, grazing boxes with tmin == tmax, tMin within a few
ulps of the entry distance, faces at the origin's coordinate, far origins,
tiny and inverted boxes).  Every decision must equal the exact one: the
traversal then visits the same nodes, so hits stay bit-exact.  The
harness's sensitivity: without the near-tie checks the fast form disagrees
on about 2% of these cases."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "ff-srl_amd"))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    import build as B
    hipcc = B._hipcc()
    if not hipcc or not os.path.exists(hipcc):
        pytest.skip("hipcc not found")
    exe = str(tmp_path_factory.mktemp("box") / "render_box_harness")
    flags = [f for f in B.HIPCC_FLAGS if f not in ("-fPIC",)]
    cmd = [hipcc, *flags, os.path.join(ROOT, "tests", "render_box_harness.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    return exe


@pytest.mark.parametrize("seed", [11, 12, 13])
def test_fast_box_decisions_equal_exact(harness, seed):
    r = subprocess.run([harness, "2000000", str(seed)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:]
    last = r.stdout.strip().splitlines()[-1].split()
    assert last[:2] == ["cases", "2000000"] and last[-2:] == ["mismatches", "0"]
