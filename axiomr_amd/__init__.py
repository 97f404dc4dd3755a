"""MI355X (gfx950) tiled render path with the reference renderer's pipeline
interface.  The compute runs in libaxiomr_amd.so (HIP kernels behind the C
ABI of include/axiomr.h); this package is the host-side mirror of the
reference's Pipeline / TiledPipeline / IShader / Mesh / Texture /
Framebuffer / Camera classes plus the synthetic workloads."""
from ._native import AxrError, NativeLibraryMissing, lib  # noqa: F401
from .render import (  # noqa: F401
    Batch, Camera, Color, Context, FlatColorShader, FlatShader, Framebuffer, GouraudShader, IShader, InputState,
    Material, MaterialGroup, Mesh, PBRShader, PhongPointShader, PhongShader, Texture, TiledPipeline, camera_matrices,
    draw_params,
)

__all__ = [
    "AxrError", "NativeLibraryMissing", "lib", "Batch", "Camera", "Color", "Context", "FlatColorShader",
    "FlatShader", "Framebuffer", "GouraudShader", "IShader", "InputState", "Material", "MaterialGroup", "Mesh", "PBRShader",
    "PhongPointShader", "PhongShader", "Texture", "TiledPipeline", "camera_matrices", "draw_params",
]
