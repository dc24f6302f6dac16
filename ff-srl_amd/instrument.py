"""Laparoscope geometry: the seven transforms of one instrument and the
capsule points the reference derives from them at construction.

The reference reads the instrument from USD prims (``laparoscope.py:302-386``):
a rotation centre (the trocar, the remote centre of motion), a rod with its
capsule, and two jaws that turn about pivots at the rod's end.  No USD reader
exists here, so ``Laparoscope`` builds the same hierarchy from a few
dimensions.  Transform order and meaning follow ``laparoscope.py:375-381``:

    0 rod capsule, 1 left jaw, 2 right jaw, 3 rotation centre,
    4 left jaw pivot, 5 right jaw pivot, 6 rod

Every transform maps local to world coordinates.  The rod's local +z points
from the rotation centre toward the jaws.  A jaw pivot is rotated about its
local x by the jaw angle, and the right pivot's angle relative to the rod is
the one the clamp state machine reads (``modelDO.py:323-324``).
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional, Tuple

import numpy as np


def _rot(axis: int, a: float) -> np.ndarray:
    c, s = math.cos(a), math.sin(a)
    m = np.eye(4)
    i, j = [(1, 2), (2, 0), (0, 1)][axis]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


def _shift(x: float, y: float, z: float) -> np.ndarray:
    m = np.eye(4)
    m[:3, 3] = (x, y, z)
    return m


def transformation(sx, sy, sz, ax, ay, az) -> np.ndarray:
    """getTransformationMatrix (``modelDO.py:396-416``): T(s) Rx Ry Rz, float64."""
    return _shift(sx, sy, sz) @ _rot(0, ax) @ _rot(1, ay) @ _rot(2, az)


def _point(m: np.ndarray, p) -> np.ndarray:
    return (m @ np.array([p[0], p[1], p[2], 1.0]))[:3]


@dataclasses.dataclass
class Laparoscope:
    """One instrument: rotation centre position and Rx/Ry/Rz angles (radians),
    the rod's insertion (rotation centre to rod centre), rod and jaw capsule
    heights and radii, and the jaw opening in degrees (40 closed .. 85 open,
    ``modelDO.py:20-22``)."""
    rcm: Tuple[float, float, float] = (0.0, 2.0, 0.0)
    rcm_rotation: Tuple[float, float, float] = (math.pi / 2, 0.0, 0.0)   # rod +z points down (world -y)
    insertion: float = 0.5
    rod_height: float = 0.6
    rod_radius: float = 0.015
    jaw_height: float = 0.06
    jaw_radius: float = 0.008
    jaw_angle: float = 60.0
    # a scene's own transforms and per-jaw capsules (from_transforms)
    xf: Optional[np.ndarray] = dataclasses.field(default=None, repr=False)
    left_jaw: Optional[Tuple[float, float]] = None     # (height, radius)
    right_jaw: Optional[Tuple[float, float]] = None

    @classmethod
    def from_transforms(cls, xf, rod_height, rod_radius, left_height, left_radius, right_height,
                        right_radius) -> "Laparoscope":
        """An instrument given by its seven local-to-world transforms (the
        order above, column-vector form) and capsule sizes, as a USD scene
        defines it."""
        xf = np.asarray(xf, np.float64).reshape(7, 4, 4)
        return cls(rod_height=rod_height, rod_radius=rod_radius, jaw_height=right_height, jaw_radius=right_radius,
                   xf=xf.copy(), left_jaw=(left_height, left_radius), right_jaw=(right_height, right_radius))

    @classmethod
    def aimed_at(cls, point, rcm_height: float = 1.0, **kw) -> "Laparoscope":
        """A vertical instrument whose drag point sits at `point`, entering
        through a rotation centre `rcm_height` above it."""
        lap = cls(**kw)
        p = np.asarray(point, np.float64)
        reach = lap.rod_height / 2.0 + 1.2 * lap.jaw_height          # rod centre to drag point
        lap.rcm = (float(p[0]), float(p[1] + rcm_height), float(p[2]))
        lap.rcm_rotation = (math.pi / 2, 0.0, 0.0)
        lap.insertion = rcm_height - reach
        return lap

    def transforms(self) -> np.ndarray:
        """(7, 4, 4) local-to-world matrices, column-vector convention, float64."""
        if self.xf is not None:
            return self.xf.copy()
        rc = transformation(*self.rcm, *self.rcm_rotation)
        rod = rc @ _shift(0.0, 0.0, self.insertion)
        end = rod @ _shift(0.0, 0.0, self.rod_height / 2.0)
        a = math.radians(self.jaw_angle)
        left_pivot = end @ _rot(0, -a)
        right_pivot = end @ _rot(0, a)
        half = _shift(0.0, 0.0, self.jaw_height / 2.0)
        return np.stack([rod, left_pivot @ half, right_pivot @ half, rc, left_pivot, right_pivot, rod])

    def initial_points(self):
        """Capsule points as the reference computes them at construction
        (``laparoscope.py:324-363, 375-386``): tips, bases (4, 3), radii (4,)
        = twice the capsule radii (0 for the drag point) and heights (4,)."""
        x = self.transforms()
        rod, left, right = x[6], x[1], x[2]
        hr = self.rod_height
        hl, rl = self.left_jaw or (self.jaw_height, self.jaw_radius)
        hrj, rr = self.right_jaw or (self.jaw_height, self.jaw_radius)
        rod_base, rod_tip = _point(rod, (0, 0, -hr / 2.0)), _point(rod, (0, 0, hr / 2.0))
        lb, lt = _point(left, (0, 0, -hl / 1.9)), _point(left, (0, 0, hl / 1.9))
        rb, rt = _point(right, (0, 0, -hrj / 1.9)), _point(right, (0, 0, hrj / 1.9))
        drag = _point(rod, (0, 0, hr / 2.0 + 1.2 * hrj))
        tips = np.stack([rod_tip, lt, rt, drag])
        bases = np.stack([rod_base, lb, rb, rod_base])
        radii = np.array([2 * self.rod_radius, 2 * rl, 2 * rr, 0.0])
        heights = np.array([hr, hl, hrj, 0.0])
        return tips, bases, radii, heights

    def stored_transforms(self) -> np.ndarray:
        """The seven transforms as the library stores them: transposed, float32."""
        return np.ascontiguousarray(np.transpose(self.transforms(), (0, 2, 1)), np.float32)
