"""Ray-cast renderer (``FF_SRL/FF_SRL/render.py:826-1196``, ``WarpRaycastRendererDO``).

The camera follows the reference: ``tan = verticalAperture / (2 * 50)``,
``aspect = horizontal / vertical``, rotation ``quat_rpy(rot[2], rot[0],
rot[1])`` (``render.py:866-868``), ``lookAt`` sets pitch = asin(d.y) and
yaw = atan2(d.x, d.z) - pi (``render.py:1186-1196``); rays sample a 2x
supersampled grid with ``sx = 2x/W - 1``, ``sy = 2y/H - 1``, ``y = H - row``
(``render.py:299-311``) and pixels are the 2x2 box mean (``render.py:699-724``).
The ray directions of the fixed camera are computed once on the host in
float64 and rounded to float32; every environment reads the same table.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import torch

from ._lib import Camera, lib
from .model import SimModelDO


class RenderMode:
    grayscale = 0
    texture = 1          # the reference's default: texture lookup, vertex colours where a mesh has no texture
    normal_map = 2
    vertex_color = 3
    depth_map = 4
    triangle_color = 5   # per-triangle colours (seeded; no reference counterpart)


def quat_rpy(roll: float, pitch: float, yaw: float) -> np.ndarray:
    """(x, y, z, w) quaternion of R = Rz(yaw) Ry(pitch) Rx(roll)."""
    cy, sy = math.cos(yaw * 0.5), math.sin(yaw * 0.5)
    cr, sr = math.cos(roll * 0.5), math.sin(roll * 0.5)
    cp, sp = math.cos(pitch * 0.5), math.sin(pitch * 0.5)
    w = cy * cr * cp + sy * sr * sp
    x = cy * sr * cp - sy * cr * sp
    y = cy * cr * sp + sy * sr * cp
    z = sy * cr * cp - cy * sr * sp
    return np.array([x, y, z, w], dtype=np.float64)


def quat_rotate(q: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Rotate rows of v (N, 3) by the unit quaternion q = (x, y, z, w)."""
    u, w = q[:3], q[3]
    return v * (2.0 * w * w - 1.0) + np.cross(u[None, :], v) * (2.0 * w) + u[None, :] * (2.0 * (v @ u))[:, None]


def ray_directions(height: int, width: int, rot_q: np.ndarray, v_aperture: float = 19.2,
                   h_aperture: float = 19.2) -> np.ndarray:
    """float32 (2H*2W, 4) unit directions of the supersampled ray grid."""
    tan = v_aperture / (2.0 * 50.0)
    aspect = h_aperture / v_aperture
    RW, RH = 2 * width, 2 * height
    r = np.arange(RW * RH)
    x = (r % RW).astype(np.float64)
    y = (RH - r // RW).astype(np.float64)
    sx = 2.0 * x / RW - 1.0
    sy = 2.0 * y / RH - 1.0
    v = np.stack([sx * tan * aspect, sy * tan, -np.ones_like(sx)], axis=1)
    d = quat_rotate(rot_q, v)
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    out = np.zeros((RW * RH, 4), dtype=np.float32)
    out[:, :3] = d.astype(np.float32)
    return out


class RaycastRendererDO:
    """Headless multi-environment ray caster.  ``renderNew`` returns
    ``(numEnvs, H, W, 3)`` like the reference (``render.py:1099-1161``);
    ``render_depth`` returns the ``(numEnvs, H, W)`` depth directly."""

    def __init__(self, device, simModel: SimModelDO, cameraPos=(3.0, 3.0, 3.0), cameraRot=(0.0, 0.0, 0.0),
                 lightPos=(3.0, 0.0, 0.0), lightIntensity=0.3, resolution: int = 512, mode: str = "headless",
                 horizontalAperture=19.2, verticalAperture=19.2, render_mode: int = RenderMode.depth_map,
                 traversal: int = 0):
        if resolution % 4:
            raise ValueError("resolution must be a multiple of 4")
        self.device = torch.device(device)
        self.simModel = simModel
        self.mode = mode
        self.height = self.width = int(resolution)
        self.render_mode = render_mode
        self.traversal = int(traversal)
        self.h_aperture, self.v_aperture = horizontalAperture, verticalAperture
        self.cameraPos = np.asarray(cameraPos, dtype=np.float64)
        self.cameraRot = np.asarray(cameraRot, dtype=np.float64)
        self.lightPos = np.asarray(lightPos, dtype=np.float32)
        self.lightIntensity = float(lightIntensity)
        n = simModel.numEnvs
        self.depth = torch.zeros((n, self.height, self.width), dtype=torch.float32, device=self.device)
        self.pixels = torch.zeros((n, self.height, self.width, 3), dtype=torch.float32, device=self.device)
        self._update_camera()

    def _update_camera(self):
        q = quat_rpy(self.cameraRot[2], self.cameraRot[0], self.cameraRot[1])
        self.rot_q = q
        self.dirs = torch.from_numpy(ray_directions(self.height, self.width, q, self.v_aperture,
                                                    self.h_aperture)).to(self.device)
        cam = Camera()
        cam.origin[:] = tuple(float(c) for c in self.cameraPos.astype(np.float32))
        cam.dirs = self.dirs.data_ptr()
        cam.height, cam.width = self.height, self.width
        cam.mode = int(self.render_mode)
        cam.light_dir[:] = tuple(float(c) for c in self.lightPos)
        cam.light_intensity = self.lightIntensity
        cam.traversal = self.traversal
        self.camera = cam

    def lookAt(self, at) -> None:
        d = np.asarray(at, dtype=np.float64) - self.cameraPos.astype(np.float32).astype(np.float64)
        d = d / np.linalg.norm(d)
        self.cameraRot[2] = math.asin(d[1])
        self.cameraRot[0] = math.atan2(d[0], d[2]) - math.pi
        self._update_camera()

    # Camera controls of the reference (render.py:1163-1184).  Position and
    # angles accumulate in float32, as the reference's wp.vec3 fields do.
    def _move(self, delta) -> None:
        self.cameraPos = (self.cameraPos.astype(np.float32) + np.asarray(delta, np.float32)).astype(np.float64)
        self.camera.origin[:] = tuple(float(c) for c in self.cameraPos.astype(np.float32))

    def moveCameraForward(self, distance: float = 0.1) -> None:
        self._move((0.0, 0.0, -distance))

    def moveCameraLeft(self, distance: float = 0.1) -> None:
        self._move((-distance, 0.0, 0.0))

    def moveCameraUp(self, distance: float = 0.1) -> None:
        self._move((0.0, distance, 0.0))

    def rotateCameraHV(self, distanceX: float = 5.0, distanceY: float = 5.0, scale: float = 0.1) -> None:
        inc = np.array([distanceX * scale * math.pi / 180.0, 0.0, distanceY * scale * math.pi / 180.0], np.float32)
        self.cameraRot = (self.cameraRot.astype(np.float32) + inc).astype(np.float64)
        self._update_camera()

    def set_mode(self, render_mode: int) -> None:
        self.render_mode = render_mode
        self.camera.mode = int(render_mode)

    def set_traversal(self, traversal: int) -> None:
        """0 picks by image size (2x2-ray quads per lane from 128x128 pixels,
        one ray per lane below), 1 one ray per lane, 2 quads.  The images are
        the same bit for bit."""
        self.traversal = int(traversal)
        self.camera.traversal = self.traversal

    def depth_camera(self) -> Camera:
        """A copy of the camera in depth mode (the renderer keeps its mode)."""
        cam = Camera()
        C.pointer(cam)[0] = self.camera
        cam.mode = RenderMode.depth_map
        return cam

    def render_depth(self, env_begin: int = 0, env_end: int = None, hit=None, thit=None) -> torch.Tensor:
        m = self.simModel
        ee = m.numEnvs if env_end is None else env_end
        cam = self.camera if self.camera.mode == RenderMode.depth_map else self.depth_camera()
        out = self.depth[env_begin:ee]
        lib.ffsrl_render(m.batch, cam, env_begin, ee, out.data_ptr(), None, None, None, None,
                         None if hit is None else hit.data_ptr(), None if thit is None else thit.data_ptr(),
                         m.stream())
        return out

    def render_rgb(self, env_begin: int = 0, env_end: int = None) -> torch.Tensor:
        m = self.simModel
        ee = m.numEnvs if env_end is None else env_end
        if self.render_mode == RenderMode.depth_map:
            raise ValueError("render_rgb needs a colour render mode")
        lib.ffsrl_vertex_normals(m.batch, env_begin, ee, m.normals.data_ptr(), m.stream())
        out = self.pixels[env_begin:ee]
        lib.ffsrl_render(m.batch, self.camera, env_begin, ee, None, out.data_ptr(), m.normals.data_ptr(),
                         m.vertexColors.data_ptr(), m.t_tri_colors.data_ptr(), None, None, m.stream())
        return out

    def render_rgbd(self, env_begin: int = 0, env_end: int = None):
        """Colour and depth images from one traversal: ``(pixels, depth)``,
        each equal to what ``render_rgb`` and ``render_depth`` return."""
        m = self.simModel
        ee = m.numEnvs if env_end is None else env_end
        if self.render_mode == RenderMode.depth_map:
            raise ValueError("render_rgbd needs a colour render mode")
        lib.ffsrl_vertex_normals(m.batch, env_begin, ee, m.normals.data_ptr(), m.stream())
        pix, dep = self.pixels[env_begin:ee], self.depth[env_begin:ee]
        lib.ffsrl_render(m.batch, self.camera, env_begin, ee, dep.data_ptr(), pix.data_ptr(), m.normals.data_ptr(),
                         m.vertexColors.data_ptr(), m.t_tri_colors.data_ptr(), None, None, m.stream())
        return pix, dep

    def renderNew(self, simBVH=None, simModel=None) -> torch.Tensor:
        """``(numEnvs, H, W, 3)`` image of every environment."""
        if self.render_mode == RenderMode.depth_map:
            return self.render_depth().unsqueeze(-1).expand(-1, -1, -1, 3)
        return self.render_rgb()

    def render(self, simBVH=None, simModel=None, useGraph: bool = False) -> torch.Tensor:
        """The reference's per-environment form (``render.py:1024-1097``): one
        launch per environment; the same images as ``renderNew``."""
        n = self.simModel.numEnvs
        for i in range(n):
            if self.render_mode == RenderMode.depth_map:
                self.render_depth(i, i + 1)
            else:
                self.render_rgb(i, i + 1)
        if self.render_mode == RenderMode.depth_map:
            return self.depth.unsqueeze(-1).expand(-1, -1, -1, 3)
        return self.pixels


# The reference's class name, for drop-in use.
WarpRaycastRendererDO = RaycastRendererDO
