"""Writes tests/golden/liver_retraction.npz: the reference's liver
retraction scene (FF_SRL/FF_SRL/scenes/liverRetraction.usd, the scene of its
tests/testLiverRetraction.py) read by this package's crate reader and
converted to plain arrays, so GPU tests and the bench can use the scene on a
machine without the reference checkout.

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


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    us = importlib.import_module("ff-srl_amd.usd_scene")
    sc = us.load_usd_scene(src)
    us.save_scene(sc, OUT)
    m = sc.meshes[0]
    print(f"wrote {OUT}: {m.name} V={m.num_vertices} T={m.num_tets} E={m.num_edges} F={m.num_tris} "
          f"pinned={sc.pinned} laparoscope={'yes' if sc.laparoscope is not None else 'no'} "
          f"({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    main()
