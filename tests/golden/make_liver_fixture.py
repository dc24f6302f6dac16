"""Writes tests/golden/liver_retraction.npz: the reference's liver
retraction scene (FF_SRL/FF_SRL/scenes/liverRetraction.usd, the scene of its
tests/testLiverRetraction.py) read by this package's crate reader and
converted to plain arrays, so GPU tests and the bench can use the scene on a
machine without the reference checkout; and tests/golden/liver_render_map.npz:
its render points and their simulation vertices (mapping:pointToVertex).

Run from the repository root in a checkout next to the reference:
    python tests/golden/make_liver_fixture.py [path/to/liverRetraction.usd]
"""
import importlib
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

SRC = "/root/reference/FF_SRL/FF_SRL/scenes/liverRetraction.usd"
OUT = os.path.join(HERE, "liver_retraction.npz")
MAP_OUT = os.path.join(HERE, "liver_render_map.npz")


def render_map(src):
    """The render mesh's points (world frame, float32) and its
    mapping:pointToVertex pairs (render point, simulation vertex): the
    inputs of remapAToBKernel (utils.py:32-42) in the reference's scene."""
    import numpy as np
    us = importlib.import_module("ff-srl_amd.usd_scene")
    uc = importlib.import_module("ff-srl_amd.usdc")
    crate = uc.CrateFile(src)
    prim = next(p for p in crate.prim_paths() if crate.get(p, "extMesh:elem") is not None)
    w = us.world_matrix(crate, prim)
    points = us._points(w, np.asarray(crate.get(prim, "points"), np.float64)).astype(np.float32)
    pairs = np.asarray(crate.get(prim, "mapping:pointToVertex"), np.int32).reshape(-1, 2)
    return points, pairs


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    us = importlib.import_module("ff-srl_amd.usd_scene")
    sc = us.load_usd_scene(src)
    us.save_scene(sc, OUT)
    points, pairs = render_map(src)
    import numpy as np
    np.savez_compressed(MAP_OUT, render_points=points, point_to_vertex=pairs)
    print(f"wrote {MAP_OUT}: {points.shape[0]} render points")
    m = sc.meshes[0]
    print(f"wrote {OUT}: {m.name} V={m.num_vertices} T={m.num_tets} E={m.num_edges} F={m.num_tris} "
          f"pinned={sc.pinned} laparoscope={'yes' if sc.laparoscope is not None else 'no'} "
          f"({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
