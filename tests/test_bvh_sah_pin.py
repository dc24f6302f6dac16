"""Pins the binned-SAH BVH build (bvh.py:281-401, 456-518) independently of
the oracle's own code.

The brute force below restates findBestSplitPlaneBinned from its
definition, without the prefix sweeps: for every axis whose centroid
bounds differ it forms the 7 candidate planes lo + (hi - lo)/8 * j, puts each
triangle in bin min(7, int((c - lo) * 8 / (hi - lo))), builds the left set
(bins < j) and the right set explicitly, and costs each plane as
|L| * area(box(L)) + |R| * area(box(R)) from the triangles' own boxes.  The
chosen plane is the first minimum (strict <, axes 0..2, planes 1..7); a node
stays a leaf when that cost is >= triCount * area(node box) (bvh.py:349-364)
or the partition c < pos leaves a side empty (:378-380); children are
allocated in pairs, left subtree first (:382-398).

Every node of the oracle's tree (and of the product's C++ host build) is
checked against it: the same node count, children, levels, and triangle set
per node.  The in-place two-pointer order (:367-376) is pinned by a
hand-worked example.

Plausible misreadings of the reference, planted in the oracle on the CPU
(``oracle.bvh_build(mutant=...)``), each fail these checks:
  * ``partition_le``            - <= instead of < in the partition;
  * ``right_sweep_off_by_one``  - the right side of plane j starting one bin
                                  late in the sweep (:329-340);
  * ``no_bin_clamp``            - no min(bins - 1, ...) on the bin index, so
                                  the top centroid falls out of the costs.
"""
import ctypes as C

import numpy as np
import pytest

BINS = 8


def _area(lo, hi):
    e = np.asarray(hi, np.float64) - np.asarray(lo, np.float64)
    return e[0] * e[1] + e[1] * e[2] + e[2] * e[0]


class Brute:
    """Binned SAH from its definition (no prefix sums)."""

    def __init__(self, vertex, faces):
        self.v = np.asarray(vertex, np.float32)
        self.f = np.asarray(faces, np.int64).reshape(-1, 3)
        tv = self.v[self.f]
        self.cen = ((tv[:, 0] + tv[:, 1]) + tv[:, 2]) / np.float32(3.0)   # float32, calculateCentroids
        self.tlo = tv.astype(np.float64).min(axis=1)
        self.thi = tv.astype(np.float64).max(axis=1)

    def box(self, S):
        S = list(S)
        return self.tlo[S].min(axis=0), self.thi[S].max(axis=0)

    def cost_of(self, S):
        if not S:
            return 0.0
        lo, hi = self.box(S)
        return len(S) * _area(lo, hi)

    def split(self, S):
        best = (1e30, -1, 0.0)
        for axis in range(3):
            c = [float(self.cen[t, axis]) for t in S]
            lo, hi = min(c), max(c)
            if lo == hi:
                continue
            for j in range(1, BINS):
                L = [t for t, ct in zip(S, c) if min(BINS - 1, int((ct - lo) * (BINS / (hi - lo)))) < j]
                R = [t for t, ct in zip(S, c) if min(BINS - 1, int((ct - lo) * (BINS / (hi - lo)))) >= j]
                cost = self.cost_of(L) + self.cost_of(R)
                if cost < best[0]:
                    best = (cost, axis, lo + (hi - lo) / BINS * j)
        return best

    def tree(self):
        """Nodes in allocation order: [triangle set, left child or -1, level]."""
        nf = self.f.shape[0]
        nodes = [[list(range(nf)), -1, 0]]
        root_box = (self.v.min(axis=0).astype(np.float64), self.v.max(axis=0).astype(np.float64))

        def sub(n, box):
            S, _, lvl = nodes[n]
            cost, axis, pos = self.split(S)
            if cost >= len(S) * _area(*box):
                return
            L = [t for t in S if float(self.cen[t, axis]) < pos]
            R = [t for t in S if not float(self.cen[t, axis]) < pos]
            if not L or not R:
                return
            a = len(nodes)
            nodes.extend([[L, -1, lvl + 1], [R, -1, lvl + 1]])
            nodes[n][1] = a
            sub(a, self.box(L))
            sub(a + 1, self.box(R))

        sub(0, root_box)
        return nodes


def _node_sets(t):
    """Triangle set of every node of a built tree (leaves from first/count
    and tri_map, inner nodes as the union of their children)."""
    n = len(t["left"])
    sets = [None] * n

    def rec(k):
        if t["left"][k] < 0:
            sets[k] = set(t["tri_map"][t["first"][k]:t["first"][k] + t["count"][k]].tolist())
        else:
            a = t["left"][k]
            sets[k] = rec(a) | rec(a + 1)
        return sets[k]

    rec(0)
    return sets


def _mismatch(vertex, faces, t):
    """None if the tree equals the brute force node for node, else why not."""
    ref = Brute(vertex, faces).tree()
    n = len(t["left"])
    if n != len(ref):
        return f"{n} nodes, brute force {len(ref)}"
    sets = _node_sets(t)
    for k, (S, left, lvl) in enumerate(ref):
        if int(t["left"][k]) != left:
            return f"node {k}: left {int(t['left'][k])}, brute force {left}"
        if int(t["level"][k]) != lvl:
            return f"node {k}: level {int(t['level'][k])}, brute force {lvl}"
        if sets[k] != set(S):
            return f"node {k}: triangle set differs"
        if (left < 0) != (int(t["count"][k]) > 0):
            return f"node {k}: leaf flag"
    return None


def _random_scenes(count=60, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nf = int(rng.integers(2, 17))
        nv = 3 * nf
        v = rng.uniform(-1.0, 1.0, size=(nv, 3)).astype(np.float32)
        if rng.uniform() < 0.5:                        # clustered, flattened, stretched variants
            v[:, int(rng.integers(3))] *= np.float32(rng.uniform(0.01, 0.3))
        f = rng.permutation(nv).reshape(nf, 3).astype(np.int32)
        out.append((v, f))
    return out


def _grid_scene(n=9):
    """n small triangles with float32-exact centroids x = 0 .. n-1 (the
    candidate planes of the root fall exactly on centroids)."""
    v, f = [], []
    for k in range(n):
        b = len(v)
        v += [(k - 0.25, 0.0, 0.0), (k + 0.25, 0.0, 0.0), (float(k), 0.5, 0.0)]
        f.append((b, b + 1, b + 2))
    return np.asarray(v, np.float32), np.asarray(f, np.int32)


def _clusters():
    """Two clusters of 4 triangles each, 10 apart along x."""
    rng = np.random.default_rng(3)
    v, f = [], []
    for cx in (0.0, 10.0):
        for _ in range(4):
            b = len(v)
            v += [tuple(np.array([cx, 0.0, 0.0]) + rng.uniform(-0.5, 0.5, 3)) for _ in range(3)]
            f.append((b, b + 1, b + 2))
    return np.asarray(v, np.float32), np.asarray(f, np.int32)


def _scenes():
    return _random_scenes() + [_grid_scene(), _grid_scene(5), _clusters()]


def test_brute_force_pins_the_oracle(orc):
    for i, (v, f) in enumerate(_scenes()):
        why = _mismatch(v, f, orc.bvh_build(v, f))
        assert why is None, f"scene {i}: {why}"


def test_brute_force_pins_the_host_build(pkg):
    """The product's C++ host builder (ffsrl_bvh_build) against the brute
    force directly."""
    for i, (v, f) in enumerate(_scenes()):
        nf = f.shape[0]
        out = {k: np.empty(2 * nf, np.int32) for k in ("left", "first", "count", "level")}
        out["tri_map"] = np.empty(nf, np.int32)
        nn, ml = C.c_int32(0), C.c_int32(0)
        vv, ff = np.ascontiguousarray(v), np.ascontiguousarray(f)
        pkg.lib.ffsrl_bvh_build(vv.ctypes.data, vv.shape[0], ff.ctypes.data, nf, BINS, out["left"].ctypes.data,
                                out["first"].ctypes.data, out["count"].ctypes.data, out["level"].ctypes.data,
                                out["tri_map"].ctypes.data, C.byref(nn), C.byref(ml))
        t = {k: a[:nn.value] for k, a in out.items() if k != "tri_map"}
        t["tri_map"] = out["tri_map"]
        why = _mismatch(v, f, t)
        assert why is None, f"scene {i}: {why}"


@pytest.mark.parametrize("mutant", ["partition_le", "right_sweep_off_by_one", "no_bin_clamp"])
def test_planted_mistakes_fail(orc, mutant):
    fails = [i for i, (v, f) in enumerate(_scenes()) if _mismatch(v, f, orc.bvh_build(v, f, mutant=mutant))]
    assert fails, f"mutant {mutant} passed every scene"


def test_two_clusters_split_between_them(orc):
    v, f = _clusters()
    t = orc.bvh_build(v, f)
    sets = _node_sets(t)
    a = int(t["left"][0])
    assert a == 1 and sets[1] == {0, 1, 2, 3} and sets[2] == {4, 5, 6, 7}


def test_equal_centroids_stay_one_leaf(orc):
    """Triangles sharing one centroid: every axis has equal bounds, no plane
    is a candidate (best cost stays 1e30), and the root is a leaf."""
    v = np.array([(-1, 0, 0), (1, 0, 0), (0, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, 0),
                  (0, 0, -1), (0, 0, 1), (0, 0, 0)], np.float32)
    f = np.arange(9, dtype=np.int32).reshape(3, 3)
    t = orc.bvh_build(v, f)
    assert len(t["left"]) == 1 and int(t["left"][0]) == -1 and int(t["count"][0]) == 3
    assert Brute(v, f).split([0, 1, 2])[0] == 1e30


def test_grid_scene_first_minimum(orc):
    """Centroids x = 0..8 with planes on centroids: the SAH costs are
    symmetric about x = 4, so the first minimum is the plane x = 4, and
    the centroid at x = 4 goes right (c < pos is false)."""
    v, f = _grid_scene()
    cost, axis, pos = Brute(v, f).split(list(range(9)))
    assert axis == 0 and pos == 4.0
    sets = _node_sets(orc.bvh_build(v, f))
    assert sets[1] == {0, 1, 2, 3} and sets[2] == {4, 5, 6, 7, 8}


def test_two_pointer_partition_hand_worked(orc):
    """bvh.py:367-376 on ids 0..5 with centroids x = 5, 1, 4, 0, 3, 2 and
    pos = 2.5, worked by hand:
        i=0 j=5  c(0)=5 -> swap 0,5: [5 1 2 3 4 0] j=4
        i=0      c(5)=2 -> i=1;  c(1)=1 -> i=2
        i=2 j=4  c(2)=4 -> swap 2,4: [5 1 4 3 2 0] j=3
        i=2 j=3  c(4)=3 -> swap 4,3: [5 1 3 4 2 0] j=2
        i=2      c(3)=0 -> i=3 > j: stop, 3 on the left."""
    cen = np.array([5.0, 1.0, 4.0, 0.0, 3.0, 2.0], np.float32)
    tri = np.arange(6, dtype=np.int64)
    assert orc.sah_partition(tri, 0, 6, cen, 2.5) == 3
    assert tri.tolist() == [5, 1, 3, 4, 2, 0]
    # a sub-range leaves the rest alone
    tri = np.arange(8, dtype=np.int64)
    cen8 = np.array([9.0, 5.0, 1.0, 4.0, 0.0, 3.0, 2.0, 9.0], np.float32)
    assert orc.sah_partition(tri, 1, 6, cen8, 2.5) == 3
    assert tri.tolist() == [0, 6, 2, 4, 5, 3, 1, 7]
