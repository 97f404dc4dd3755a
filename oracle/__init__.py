"""ORACLE - test infrastructure only.

ctypes binding of oracle/_build/liboracle.so, the CPU restatement of the
reference renderer's tiled pipeline (reference: src/tiled_pipeline.cpp,
src/pipeline.cpp, include/shaders/shaders.hpp; see oracle/src/*.hpp for the
line-by-line map).  Only the test suite, the smoke check and the CPU-baseline
leg of bench.py use it, and only as the checker or the timed baseline; the
product library never loads it.

Pinned by (tests/test_oracle_*.py): closed forms of the setup and edge
functions, clipper invariants and hand-solved cases, brute-force per-pixel
argmin of the depth key on small scenes, bilinear sampling against a hand
computation, shading against float64 closed forms, and the golden vectors
under tests/golden/.  The GLM evaluation orders it adopts are not pinned
against GLM itself (GLM is not vendored in the reference).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

_F = C.POINTER(C.c_float)
_U32 = C.POINTER(C.c_uint32)
_U8 = C.POINTER(C.c_uint8)
_P = C.c_void_p


class OrcDrawParams(C.Structure):
    _fields_ = [("model", C.c_float * 16), ("view_proj", C.c_float * 16), ("viewport", C.c_float * 16),
                ("camera_pos", C.c_float * 3), ("shader", C.c_int32), ("light_dir", C.c_float * 3),
                ("light_color", C.c_float * 3), ("num_lights", C.c_int32), ("light_pos", (C.c_float * 3) * 4),
                ("light_colors", (C.c_float * 3) * 4)]


class OrcMaterial(C.Structure):
    _fields_ = [("diffuse", _P), ("metallic", _P), ("bump", _P), ("roughness", _P), ("ao", _P),
                ("specular_exponent", C.c_float)]


_SIGS = {
    "orc_texture_create": (_P, [_U8, C.c_int, C.c_int, C.c_int]),
    "orc_texture_destroy": (None, [_P]),
    "orc_mesh_create": (_P, [_F, C.c_size_t, _U32, C.c_size_t, _U32, C.c_int, C.POINTER(OrcMaterial), C.c_int, _U32]),
    "orc_mesh_destroy": (None, [_P]),
    "orc_fb_create": (_P, [C.c_int, C.c_int]),
    "orc_fb_destroy": (None, [_P]),
    "orc_fb_clear": (None, [_P, C.c_uint32, C.c_float]),
    "orc_fb_read": (None, [_P, _U8, _F, _U32]),
    "orc_pipeline_create": (_P, [C.c_int]),
    "orc_pipeline_destroy": (None, [_P]),
    "orc_draw": (C.c_int, [_P, _P, _P, C.POINTER(OrcDrawParams)]),
    "orc_phase_times": (None, [_P, C.POINTER(C.c_double)]),
    "orc_last_triangle_count": (C.c_size_t, [_P]),
    "orc_clip_triangle": (C.c_int, [_F, _F]),
    "orc_clip_single_plane": (C.c_int, [C.c_int, _F, _F]),
    "orc_setup_triangle": (C.c_int, [_F, C.c_int, C.c_int, _F]),
    "orc_texture_sample": (None, [_P, C.c_float, C.c_float, _F]),
    "orc_mat4_mul": (None, [_F, _F, _F]),
    "orc_mat4_vec4": (None, [_F, _F, _F]),
}

_lib = None


REFERENCE = Path("/root/reference")
STB_PATH = HERE / "_ref" / "libstb_image.so"


def build() -> Path:
    """Compile the restatement with its Makefile (g++ -O3 -mavx2 -ffp-contract=off),
    and, where the reference tree is present, its stb_image into _ref/."""
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if (REFERENCE / "external" / "stb_image" / "stb_image.cpp").exists():
        r = subprocess.run(["make", "-s", "-C", str(HERE), "ref", f"REFERENCE={REFERENCE}"], capture_output=True,
                           text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference stb_image build failed:\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


_stb = None


def stb_image():
    """The reference's stb_image built from its source (oracle/_ref), or None
    where it has not been built."""
    global _stb
    if _stb is None:
        if not STB_PATH.exists():
            return None
        h = C.CDLL(str(STB_PATH))
        h.stbi_load_from_memory.restype = C.POINTER(C.c_uint8)
        h.stbi_load_from_memory.argtypes = [_U8, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.POINTER(C.c_int), C.c_int]
        h.stbi_image_free.restype = None
        h.stbi_image_free.argtypes = [C.c_void_p]
        _stb = h
    return _stb


def stbi_load_rgba(data: bytes):
    """stbi_load_from_memory(..., 4) of the reference's decoder: (h, w, 4)
    uint8 array, or None if it rejects the file."""
    h = stb_image()
    buf = np.frombuffer(data, np.uint8).copy()
    w, hh, n = C.c_int(), C.c_int(), C.c_int()
    p = h.stbi_load_from_memory(buf.ctypes.data_as(_U8), len(data), C.byref(w), C.byref(hh), C.byref(n), 4)
    if not p:
        return None
    out = np.ctypeslib.as_array(p, shape=(hh.value, w.value, 4)).copy()
    h.stbi_image_free(p)
    return out


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _f(a):
    return a.ctypes.data_as(_F)


class OracleScene:
    """The restated pipeline's copies of one SceneData's mesh and textures."""

    def __init__(self, scene):
        L = lib()
        self._tex = {}
        for name, t in scene.textures.items():
            rgba = np.ascontiguousarray(t.rgba, np.uint8)
            self._tex[name] = L.orc_texture_create(rgba.ctypes.data_as(_U8), rgba.shape[1], rgba.shape[0], t.filter)
        mats = (OrcMaterial * len(scene.materials))()
        for i, m in enumerate(scene.materials):
            mats[i] = OrcMaterial(*(self._tex[n] if n else None for n in (m.diffuse, m.metallic, m.bump, m.roughness,
                                                                         m.ao)), m.specular_exponent)
        groups = np.ascontiguousarray(np.array(scene.groups, np.uint32).reshape(-1))
        v = np.ascontiguousarray(scene.vertices, np.float32)
        f = np.ascontiguousarray(scene.faces, np.uint32)
        fc = None if scene.face_colors is None else np.ascontiguousarray(scene.face_colors, np.uint32)
        self.mesh = L.orc_mesh_create(_f(v), v.shape[0], f.ctypes.data_as(_U32), f.shape[0],
                                      groups.ctypes.data_as(_U32), len(scene.groups), mats, len(scene.materials),
                                      fc.ctypes.data_as(_U32) if fc is not None else None)

    def close(self):
        L = lib()
        if self.mesh:
            L.orc_mesh_destroy(self.mesh)
            self.mesh = None
        for t in self._tex.values():
            L.orc_texture_destroy(t)
        self._tex = {}


class OracleRenderer:
    """CPU tiled pipeline with `threads` workers (reference ThreadPool)."""

    def __init__(self, threads: int = 0):
        # at most 16 workers: the host share the benchmark machines give one
        # GPU, whatever os.cpu_count() reports there
        self.threads = threads or min(16, os.cpu_count() or 1)
        self.pipe = lib().orc_pipeline_create(self.threads)

    def close(self):
        if self.pipe:
            lib().orc_pipeline_destroy(self.pipe)
            self.pipe = None

    def framebuffer(self, w, h):
        return lib().orc_fb_create(w, h)

    def draw(self, fb, oscene: OracleScene, params) -> None:
        p = OrcDrawParams()
        C.memmove(C.byref(p), C.byref(params), C.sizeof(OrcDrawParams))
        code = lib().orc_draw(self.pipe, fb, oscene.mesh, C.byref(p))
        if code != 0:
            raise RuntimeError(f"orc_draw failed with {code}")

    def phase_times(self):
        out = (C.c_double * 6)()
        lib().orc_phase_times(self.pipe, out)
        return dict(zip(["ClippingAllTriangles", "InitTiles", "BinAllTriangles", "TilePreparation",
                         "RasterizingTiles", "MergingAllTiles"], list(out)))


def fb_clear(fb, bgra: int, depth: float = float("inf")):
    lib().orc_fb_clear(fb, bgra, depth)


def fb_read(fb, w, h):
    bgra = np.empty((h, w, 4), np.uint8)
    depth = np.empty((h, w), np.float32)
    ids = np.empty((h, w), np.uint32)
    lib().orc_fb_read(fb, bgra.ctypes.data_as(_U8), _f(depth), ids.ctypes.data_as(_U32))
    return bgra, depth, ids


def fb_destroy(fb):
    lib().orc_fb_destroy(fb)
