"""MI355X-native FF-SRL per-step simulation path.

Python host code on PyTorch-ROCm calling hand-written gfx950 HIP kernels
through the C ABI of ``include/ffsrl.h``.  The directory name is not a Python
identifier; import it with ``importlib.import_module("ff-srl_amd")``.
"""
from ._lib import FFSRLError, lib
from .bvh import SimBVH
from .env import SimEnvironment
from .instrument import Laparoscope
from .integrator import SimIntegratorDO
from .model import SimModelDO
from .render import RaycastRendererDO, RenderMode, WarpRaycastRendererDO
from . import autotune
from . import scenes
from . import serve

__all__ = [
    "FFSRLError", "lib", "Laparoscope", "SimBVH", "SimIntegratorDO", "SimModelDO", "RaycastRendererDO", "RenderMode",
    "SimEnvironment", "WarpRaycastRendererDO", "autotune", "scenes", "serve",
]
