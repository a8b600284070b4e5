"""MI355X-native Voxel ReSTIR: brick-grid DDA, RIS light sampling, temporal
and spatial reservoir reuse and shading as gfx950 HIP kernels behind a C ABI
(include/vr_api.h), with the reference's World host API (include/vr_world.h)
in C++ on top.

The Python package only binds the shared library (smallproject_amd/lib/
libvoxelrestir.so); every computation runs in it.
"""
from ._lib import LIBPATH, VrConfig, VrError, VrLight, VrRenderParams, lib  # noqa: F401
from .api import Batch, Renderer, World, camera_path, default_config, focal_for_fov, log_lut, terrain_height  # noqa: F401
