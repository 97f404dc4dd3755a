"""Host-side mirror of the reference renderer's pipeline interface.

Names follow the reference (include/pipeline.hpp, include/tiled_pipeline.hpp,
include/IShader.hpp, include/shaders/shaders.hpp, include/mesh.hpp,
include/texture.hpp, include/framebuffer.hpp, include/camera.hpp) so that a
caller of ``TiledPipeline::drawMesh`` finds the same calls here:

    ctx = Context()
    fb = Framebuffer(ctx, 800, 600)
    cam = Camera((0, 0, 3), (0, 0, 0), 60.0, 800 / 600)
    cam.setViewport(0, 0, 800, 600); cam.update(0.0)
    pipe = TiledPipeline(cam, fb)
    pipe.setShader(PhongShader(lightDirection=(0, -1, 0), lightColor=(0.6, 0.6, 0.6)))
    fb.clearColor(Color(0, 0, 0, 255)); fb.clearDepth()
    pipe.drawMesh(model, mesh)

Every draw runs the HIP kernels of libaxiomr_amd.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _native as N

SHADER_FLAT, SHADER_PHONG, SHADER_PBR, SHADER_GOURAUD, SHADER_PHONG_POINT, SHADER_FLAT_COLOR = range(6)
FILTER_NEAREST, FILTER_BILINEAR = 0, 1


@dataclass
class Color:
    """reference include/math.hpp:17-19."""
    r: int
    g: int
    b: int
    a: int = 255

    def packed_bgra(self) -> int:
        # reference src/framebuffer.cpp:26-36: a<<24 | r<<16 | g<<8 | b
        return (self.a << 24) | (self.r << 16) | (self.g << 8) | self.b


class Context:
    """One GPU, one HIP stream, one frame arena, an optional HBM budget."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._lib = N.lib()
        h = C.c_void_p()
        code = self._lib.axr_ctx_create(device, hbm_budget, C.byref(h))
        if code != N.AXR_OK:
            raise N.AxrError(code, f"axr_ctx_create(device={device})")
        self.handle = h

    def close(self) -> None:
        if self.handle:
            self._lib.axr_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, code: int) -> None:
        N.check(code, self.handle)

    def sync(self) -> None:
        self.check(self._lib.axr_sync(self.handle))

    def config(self) -> dict:
        """Tunables of the binning pipeline (see axr_config in include/axiomr.h)."""
        cfg = N.AxrConfig()
        self.check(self._lib.axr_ctx_get_config(self.handle, C.byref(cfg)))
        return {"tile_size": cfg.tile_size, "raster_threads": cfg.raster_threads, "bin_capacity": cfg.bin_capacity,
                "small_footprint": cfg.small_footprint, "raster_mode": cfg.raster_mode}

    def set_config(self, **kw) -> None:
        cfg = N.AxrConfig()
        self.check(self._lib.axr_ctx_get_config(self.handle, C.byref(cfg)))
        for k, v in kw.items():
            if not hasattr(cfg, k) or k == "reserved":
                raise KeyError(k)
            setattr(cfg, k, int(v))
        self.check(self._lib.axr_ctx_set_config(self.handle, C.byref(cfg)))

    def set_tile_size(self, tile: int) -> None:
        self.set_config(tile_size=tile)

    def tile_size(self) -> int:
        return self.config()["tile_size"]

    def set_profiling(self, on: bool) -> None:
        self.check(self._lib.axr_ctx_set_profiling(self.handle, 1 if on else 0))

    def stats(self) -> N.AxrKernelStats:
        s = N.AxrKernelStats()
        self.check(self._lib.axr_ctx_get_stats(self.handle, C.byref(s)))
        return s

    def memory(self) -> tuple[int, int]:
        a, b = C.c_uint64(), C.c_uint64()
        self.check(self._lib.axr_ctx_memory(self.handle, C.byref(a), C.byref(b)))
        return a.value, b.value


class Texture:
    """reference include/texture.hpp: RGBA8 texels, row 0 first; `filter`
    selects the reference's nearest sampling or the added bilinear filter."""

    def __init__(self, ctx: Context, rgba: np.ndarray, filter: int = FILTER_NEAREST):
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        if rgba.ndim != 3 or rgba.shape[2] != 4:
            raise ValueError("texture must be H x W x 4 uint8")
        self.ctx = ctx
        self.height, self.width = rgba.shape[:2]
        self.filter = filter
        h = C.c_void_p()
        ctx.check(ctx._lib.axr_texture_create(ctx.handle, N.u8ptr(rgba), self.width, self.height, filter, C.byref(h)))
        self.handle = h

    @classmethod
    def from_file(cls, ctx: Context, path: str, filter: int = FILTER_NEAREST) -> "Texture":
        """reference Texture(const std::string& filePath) (src/texture.cpp:6-35)."""
        from .ingest import load_image
        return cls(ctx, load_image(path), filter)

    def close(self):
        if self.handle:
            self.ctx._lib.axr_texture_destroy(self.handle)
            self.handle = None


@dataclass
class Material:
    """reference include/mesh.hpp:20-34."""
    diffuseTexture: Optional[Texture] = None
    metallicTexture: Optional[Texture] = None
    bumpTexture: Optional[Texture] = None
    roughnessTexture: Optional[Texture] = None
    aoTexture: Optional[Texture] = None
    specularExponent: float = 1.0


@dataclass
class MaterialGroup:
    """reference include/mesh.hpp:35-39 (material by index instead of name)."""
    material: int
    startIndex: int
    faceCount: int


class Mesh:
    """reference include/mesh.hpp:63-101, built from arrays instead of an OBJ:
    vertices (N, 14) float32 = position, uv, normal, tangent, bitangent;
    faces (F, 3) uint32; optional per-face RGBA8 colours."""

    def __init__(self, ctx: Context, vertices: np.ndarray, faces: np.ndarray,
                 groups: Sequence[MaterialGroup], materials: Sequence[Material],
                 face_colors: Optional[np.ndarray] = None):
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 14)
        f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        self.ctx = ctx
        self.nverts, self.nfaces = v.shape[0], f.shape[0]
        gs = (N.AxrGroup * len(groups))(*[N.AxrGroup(g.material, g.startIndex, g.faceCount) for g in groups])
        ms = (N.AxrMaterial * len(materials))()
        for i, m in enumerate(materials):
            ms[i] = N.AxrMaterial(*(t.handle if t is not None else None for t in
                                    (m.diffuseTexture, m.metallicTexture, m.bumpTexture, m.roughnessTexture,
                                     m.aoTexture)), m.specularExponent)
        fc = None
        if face_colors is not None:
            fc = np.ascontiguousarray(face_colors, dtype=np.uint32)
        h = C.c_void_p()
        ctx.check(ctx._lib.axr_mesh_create(ctx.handle, N.fptr(v), self.nverts, N.u32ptr(f), self.nfaces, gs,
                                           len(groups), ms, len(materials),
                                           N.u32ptr(fc) if fc is not None else None, C.byref(h)))
        self.handle = h
        self.materials = list(materials)   # keep textures alive

    @classmethod
    def from_obj(cls, ctx: Context, path: str, filter: int = FILTER_NEAREST) -> "Mesh":
        """reference Mesh(const std::string& filePath) (src/mesh.cpp:8-27):
        the OBJ, its <dir>/<basename>.mtl materials and their textures.  A
        group whose material the .mtl lacks raises, as the reference's
        Mesh::getMaterial (.at()) does when drawn; a model without usemtl has
        no group and draws nothing."""
        from .ingest import load_obj
        model = load_obj(path)
        for g in model.groups:
            if g.material is None:
                raise N.AxrError(N.AXR_ERR_INVALID, f"material '{g.materialName}' is not in the .mtl file")
        cache = {}

        def tex(p):
            if not p:
                return None
            if p not in cache:
                cache[p] = Texture.from_file(ctx, p, filter)
            return cache[p]

        mats = [Material(tex(m.diffuseMap), tex(m.metallicMap), tex(m.bumpMap), tex(m.roughnessMap),
                         tex(m.aoMap), m.specularExponent) for m in model.materials] or [Material()]
        groups = [MaterialGroup(g.material, g.startIndex, g.faceCount) for g in model.groups] or \
            [MaterialGroup(0, 0, 0)]
        mesh = cls(ctx, model.vertices, model.faces, groups, mats)
        mesh.model = model
        mesh.textures = list(cache.values())
        return mesh

    def close(self):
        if self.handle:
            self.ctx._lib.axr_mesh_destroy(self.handle)
            self.handle = None
        for t in getattr(self, "textures", []):
            t.close()


class Framebuffer:
    """reference include/framebuffer.hpp: BGRA8 colour + FP32 depth, row 0 at
    the bottom of the image."""

    def __init__(self, ctx: Context, width: int, height: int, with_ids: bool = False):
        self.ctx = ctx
        self.width, self.height = width, height
        self.with_ids = with_ids
        h = C.c_void_p()
        ctx.check(ctx._lib.axr_fb_create(ctx.handle, width, height, 1 if with_ids else 0, C.byref(h)))
        self.handle = h
        self._color = Color(0, 0, 0, 0)

    def getWidth(self) -> int:
        return self.width

    def getHeight(self) -> int:
        return self.height

    def clearColor(self, color: Color) -> None:
        self._color = color
        self.ctx.check(self.ctx._lib.axr_fb_clear(self.handle, color.packed_bgra(), np.float32(np.inf)))

    def clearDepth(self, depth: float = float("inf")) -> None:
        # the C ABI clears both planes at once; keep the last colour
        self.ctx.check(self.ctx._lib.axr_fb_clear(self.handle, self._color.packed_bgra(), depth))

    def clear(self, color: Color, depth: float = float("inf")) -> None:
        self._color = color
        self.ctx.check(self.ctx._lib.axr_fb_clear(self.handle, color.packed_bgra(), depth))

    def read(self):
        """(bgra uint8 (H, W, 4), depth float32 (H, W), ids uint32 (H, W) or None)."""
        bgra = np.empty((self.height, self.width, 4), np.uint8)
        depth = np.empty((self.height, self.width), np.float32)
        ids = np.empty((self.height, self.width), np.uint32) if self.with_ids else None
        self.ctx.check(self.ctx._lib.axr_fb_read(self.handle, N.u8ptr(bgra), N.fptr(depth),
                                                 N.u32ptr(ids) if ids is not None else None))
        return bgra, depth, ids

    def write_png(self, path: str) -> None:
        """The colour plane as an RGBA PNG, top row first."""
        import os
        self.ctx.check(self.ctx._lib.axr_fb_write_png(self.handle, os.fsencode(path)))

    def close(self):
        if self.handle:
            self.ctx._lib.axr_fb_destroy(self.handle)
            self.handle = None


def camera_matrices(position, target, fov, aspect, near=0.1, far=100.0, vp_w=800, vp_h=600):
    """reference Camera matrices (src/camera.cpp:30-81, 191-205), column-major
    float32 arrays: view, projection, view-projection, viewport."""
    out = [np.zeros(16, np.float32) for _ in range(4)]
    pos = np.asarray(position, np.float32)
    tgt = np.asarray(target, np.float32)
    N.check(N.lib().axr_camera_matrices(N.fptr(pos), N.fptr(tgt), fov, aspect, near, far, vp_w, vp_h,
                                        *(N.fptr(o) for o in out)))
    return out


@dataclass
class InputState:
    """One frame of what the reference's Window tells Camera::handleInput
    (src/camera.cpp:100-127): mouse position, right button, W/S/A/D held.
    Any object with the same three methods (a real window) works as well."""
    mouse: tuple = (0.0, 0.0)
    rightButton: bool = False
    keys: str = ""

    def isMouseButtonDown(self, button: int) -> bool:
        return bool(self.rightButton) if button == 1 else False

    def getMousePos(self):
        return self.mouse

    def isKeyDown(self, key: str) -> bool:
        return key in self.keys


def _input_of(window) -> N.AxrCameraInput:
    mx, my = window.getMousePos()
    keys = 0
    for ch, bit in (("W", N.KEY_W), ("S", N.KEY_S), ("A", N.KEY_A), ("D", N.KEY_D)):
        if window.isKeyDown(ch):
            keys |= bit
    return N.AxrCameraInput(float(mx), float(my), 1 if window.isMouseButtonDown(1) else 0, keys)


class Camera:
    """reference include/camera.hpp / src/camera.cpp: a fly camera whose
    state lives in a native axr_camera record.  As in the reference, the view
    is identity until the first update(), and setFov/setAspectRatio do not
    rebuild the projection."""

    def __init__(self, position=(0.0, 0.0, 0.0), target=(0.0, 0.0, -1.0), fov: float = 60.0,
                 aspectRatio: float = 16.0 / 9.0):
        self._c = N.AxrCamera()
        pos = np.asarray(position, np.float32)
        tgt = np.asarray(target, np.float32)
        N.check(N.lib().axr_camera_init(C.byref(self._c), N.fptr(pos), N.fptr(tgt), float(fov),
                                        float(aspectRatio)))

    @property
    def state(self) -> N.AxrCamera:
        return self._c

    def setPosition(self, position):
        N.check(N.lib().axr_camera_set_position(C.byref(self._c), N.fptr(np.asarray(position, np.float32))))

    def setTarget(self, target):
        N.check(N.lib().axr_camera_set_target(C.byref(self._c), N.fptr(np.asarray(target, np.float32))))

    def setFov(self, fov: float):
        N.check(N.lib().axr_camera_set_fov(C.byref(self._c), float(fov)))

    def setAspectRatio(self, aspectRatio: float):
        N.check(N.lib().axr_camera_set_aspect_ratio(C.byref(self._c), float(aspectRatio)))

    def setViewport(self, x, y, w, h):
        N.check(N.lib().axr_camera_set_viewport(C.byref(self._c), int(x), int(y), int(w), int(h)))

    def handleInput(self, window):
        inp = _input_of(window)
        N.check(N.lib().axr_camera_handle_input(C.byref(self._c), C.byref(inp)))

    def update(self, deltaTime: float = 0.0):
        N.check(N.lib().axr_camera_update(C.byref(self._c), float(deltaTime)))

    def getViewMatrix(self):
        return np.array(self._c.view, np.float32)

    def getProjectionMatrix(self):
        return np.array(self._c.proj, np.float32)

    def getViewProjectionMatrix(self):
        return np.array(self._c.view_proj, np.float32)

    def getViewportMatrix(self):
        out = np.zeros(16, np.float32)
        N.check(N.lib().axr_camera_viewport_matrix(C.byref(self._c), N.fptr(out)))
        return out

    def getPosition(self):
        return np.array(self._c.position, np.float32)

    def vectors(self):
        """(forward, right, up) as float32 arrays."""
        return (np.array(self._c.forward, np.float32), np.array(self._c.right, np.float32),
                np.array(self._c.up, np.float32))

    def path(self, inputs, dts):
        """Run the reference frame loop (src/renderer.cpp:122-123: handleInput,
        then update(dt)) over recorded inputs; returns the (n, 16) view-projection
        matrices and (n, 3) positions of every frame."""
        n = len(inputs)
        arr = (N.AxrCameraInput * max(n, 1))(*[_input_of(w) for w in inputs])
        dt = np.ascontiguousarray(dts, np.float32)
        if dt.shape != (n,):
            raise ValueError("one dt per input")
        vp = np.zeros((n, 16), np.float32)
        pos = np.zeros((n, 3), np.float32)
        N.check(N.lib().axr_camera_path(C.byref(self._c), arr, N.fptr(dt), n, N.fptr(vp), N.fptr(pos)))
        return vp, pos


# ---- shaders (reference include/shaders/shaders.hpp) -------------------------

class IShader:
    kind = -1

    def fill(self, p) -> None:
        pass


@dataclass
class FlatShader(IShader):
    lightDirection: tuple = (0.0, -1.0, 0.0)
    kind = SHADER_FLAT

    def fill(self, p):
        p.light_dir[:] = [float(x) for x in self.lightDirection]


@dataclass
class PhongShader(IShader):
    lightDirection: tuple = (0.0, -1.0, 0.0)
    lightColor: tuple = (0.6, 0.6, 0.6)
    kind = SHADER_PHONG

    def fill(self, p):
        p.light_dir[:] = [float(x) for x in self.lightDirection]
        p.light_color[:] = [float(x) for x in self.lightColor]


@dataclass
class PBRShader(PhongShader):
    lightColor: tuple = (5.6, 5.6, 5.6)
    kind = SHADER_PBR


@dataclass
class GouraudShader(IShader):
    lightDirection: tuple = (0.0, -1.0, 0.0)
    kind = SHADER_GOURAUD

    def fill(self, p):
        p.light_dir[:] = [float(x) for x in self.lightDirection]


@dataclass
class PhongPointShader(IShader):
    lights: list = field(default_factory=list)   # [(position xyz, colour rgb)], up to 4
    kind = SHADER_PHONG_POINT

    def fill(self, p):
        if len(self.lights) > 4:
            raise ValueError("at most 4 point lights")
        p.num_lights = len(self.lights)
        for i, (pos, col) in enumerate(self.lights):
            p.light_pos[i][:] = [float(x) for x in pos]
            p.light_colors[i][:] = [float(x) for x in col]


@dataclass
class FlatColorShader(IShader):
    kind = SHADER_FLAT_COLOR


def draw_params(model, camera: Camera, shader: IShader, flags: int = 0) -> N.AxrDrawParams:
    """The uniforms drawMesh copies into the shader (reference
    src/tiled_pipeline.cpp:148-155)."""
    p = N.AxrDrawParams()
    p.model[:] = [float(x) for x in np.asarray(model, np.float32).reshape(16)]
    p.view_proj[:] = [float(x) for x in camera.getViewProjectionMatrix()]
    p.viewport[:] = [float(x) for x in camera.getViewportMatrix()]
    p.camera_pos[:] = [float(x) for x in camera.getPosition()]
    p.shader = shader.kind
    shader.fill(p)
    p.flags = flags
    return p


class TiledPipeline:
    """reference TiledPipeline (include/tiled_pipeline.hpp:58-62): drawMesh on
    the GPU.  The thread count argument of the reference has no meaning here."""

    def __init__(self, camera: Camera, framebuffer: Framebuffer):
        self.m_Camera = camera
        self.m_Framebuffer = framebuffer
        self.m_Shader: Optional[IShader] = None

    def setShader(self, shader: IShader):
        self.m_Shader = shader

    def setCamera(self, camera: Camera):
        self.m_Camera = camera

    def setFramebuffer(self, fb: Framebuffer):
        self.m_Framebuffer = fb

    def getViewportMat(self):
        return self.m_Camera.getViewportMatrix() if self.m_Camera else None

    def drawMesh(self, model, mesh: Mesh, asynchronous: bool = False) -> None:
        # reference tiled_pipeline.cpp:146: silently does nothing without a
        # shader, camera or framebuffer
        if self.m_Shader is None or self.m_Camera is None or self.m_Framebuffer is None:
            return
        fb = self.m_Framebuffer
        p = draw_params(model, self.m_Camera, self.m_Shader, N.AXR_DRAW_ASYNC if asynchronous else 0)
        fb.ctx.check(fb.ctx._lib.axr_draw(fb.ctx.handle, fb.handle, mesh.handle, C.byref(p)))


class Batch:
    """Many views (framebuffer, mesh, uniforms) drawn by one graph-captured
    launch sequence."""

    def __init__(self, ctx: Context, fbs: Sequence[Framebuffer], meshes: Sequence[Mesh],
                 params: Sequence[N.AxrDrawParams]):
        n = len(fbs)
        if not (n == len(meshes) == len(params)) or n == 0:
            raise ValueError("batch needs equal, non-empty lists")
        self.ctx = ctx
        fa = (C.c_void_p * n)(*[f.handle.value for f in fbs])
        ma = (C.c_void_p * n)(*[m.handle.value for m in meshes])
        pa = (N.AxrDrawParams * n)(*params)
        h = C.c_void_p()
        ctx.check(ctx._lib.axr_batch_create(ctx.handle, n, C.cast(fa, N._PP), C.cast(ma, N._PP), pa, C.byref(h)))
        self.handle = h
        self._keep = (list(fbs), list(meshes))

    def set_clear(self, color: Optional["Color"], depth: float = float("inf")) -> None:
        """Clear every view's framebuffer at the start of each launch (None: don't)."""
        if color is None:
            self.ctx.check(self.ctx._lib.axr_batch_set_clear(self.handle, 0, 0, depth))
        else:
            self.ctx.check(self.ctx._lib.axr_batch_set_clear(self.handle, 1, color.packed_bgra(), depth))

    def launch(self) -> None:
        self.ctx.check(self.ctx._lib.axr_batch_launch(self.handle))

    def close(self):
        if self.handle:
            self.ctx._lib.axr_batch_destroy(self.handle)
            self.handle = None
