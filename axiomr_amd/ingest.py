"""Host-side ingestion and image output (libaxiomr_amd.so, no GPU needed).

    model = load_obj("scene.obj")          # reference Mesh(const std::string&)
    rgba = load_image("albedo.png")        # reference Texture(const std::string&)
    write_png("frame.png", rgba)

`load_obj` follows the reference loader (src/mesh.cpp:222-415): OBJ +
<dir>/<basename>.mtl, vertices deduplicated on (position, uv, normal),
polygons fanned, tangents generated per vertex.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _native as N


@dataclass
class ObjMaterial:
    """reference Material (include/mesh.hpp:20-34), texture paths resolved
    against the .mtl's directory ("" when absent)."""
    name: str
    ambient: tuple
    diffuse: tuple
    specular: tuple
    specularExponent: float
    refractiveIndex: float
    dissolve: float
    illuminationModel: int
    diffuseMap: str = ""
    metallicMap: str = ""
    bumpMap: str = ""
    roughnessMap: str = ""
    aoMap: str = ""


@dataclass
class ObjGroup:
    """reference MaterialGroup (include/mesh.hpp:35-39) plus the index of its
    material in ObjModel.materials (None if the .mtl lacks it)."""
    materialName: str
    startIndex: int
    faceCount: int
    material: Optional[int]


@dataclass
class ObjModel:
    vertices: np.ndarray          # (N, 14) float32: position, uv, normal, tangent, bitangent
    faces: np.ndarray             # (F, 3) uint32
    groups: List[ObjGroup] = field(default_factory=list)
    materials: List[ObjMaterial] = field(default_factory=list)


def _err(lib, fn_name):
    raw = getattr(lib, fn_name)()
    return raw.decode() if raw else ""


def load_obj(path: str) -> ObjModel:
    lib = N.lib()
    h = C.c_void_p()
    code = lib.axr_obj_load(os.fsencode(path), C.byref(h))
    if code != N.AXR_OK:
        raise N.AxrError(code, _err(lib, "axr_obj_last_error"))
    try:
        nv, nf = C.c_uint64(), C.c_uint64()
        ng, nm = C.c_uint32(), C.c_uint32()
        lib.axr_obj_counts(h, C.byref(nv), C.byref(nf), C.byref(ng), C.byref(nm))
        v = np.empty((nv.value, 14), np.float32)
        f = np.empty((nf.value, 3), np.uint32)
        lib.axr_obj_arrays(h, N.fptr(v), N.u32ptr(f))
        mats = []
        for i in range(nm.value):
            m = N.AxrObjMaterial()
            lib.axr_obj_get_material(h, i, C.byref(m))
            mats.append(ObjMaterial(m.name.decode(), tuple(m.ambient), tuple(m.diffuse), tuple(m.specular),
                                    m.specular_exponent, m.refractive_index, m.dissolve, m.illumination_model,
                                    m.diffuse_map.decode(), m.metallic_map.decode(), m.bump_map.decode(),
                                    m.roughness_map.decode(), m.ao_map.decode()))
        groups = []
        for i in range(ng.value):
            g = N.AxrObjGroup()
            lib.axr_obj_get_group(h, i, C.byref(g))
            groups.append(ObjGroup(g.material.decode(), g.first_face, g.face_count,
                                   None if g.material_index == 0xFFFFFFFF else g.material_index))
        return ObjModel(v, f, groups, mats)
    finally:
        lib.axr_obj_destroy(h)


def load_image(path: str) -> np.ndarray:
    """(H, W, 4) uint8, top row first."""
    lib = N.lib()
    w, h = C.c_int(), C.c_int()
    ptr = N._U8()
    code = lib.axr_image_load(os.fsencode(path), C.byref(w), C.byref(h), C.byref(ptr))
    if code != N.AXR_OK:
        raise N.AxrError(code, _err(lib, "axr_image_last_error"))
    try:
        return np.ctypeslib.as_array(ptr, shape=(h.value, w.value, 4)).copy()
    finally:
        lib.axr_image_free(ptr)


def write_png(path: str, rgba: np.ndarray, flip_rows: bool = False) -> None:
    lib = N.lib()
    a = np.ascontiguousarray(rgba, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] != 4:
        raise ValueError("image must be H x W x 4 uint8")
    code = lib.axr_image_write_png(os.fsencode(path), a.shape[1], a.shape[0], N.u8ptr(a), 1 if flip_rows else 0)
    if code != N.AXR_OK:
        raise N.AxrError(code, _err(lib, "axr_image_last_error"))
