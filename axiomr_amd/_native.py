"""ctypes binding of ``libaxiomr_amd.so`` (the C ABI in include/axiomr.h).

The library is the only compute path: if it is missing, ``lib()`` raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libaxiomr_amd.so"

AXR_OK = 0
AXR_ERR_INVALID = -1
AXR_ERR_OOM = -2
AXR_ERR_HIP = -3
AXR_ERR_OVERFLOW = -4
AXR_ERR_BUDGET = -5
AXR_ERR_NO_DEVICE = -6
AXR_ERR_NOT_READY = -7

STATUS_NAMES = {
    AXR_OK: "AXR_OK", AXR_ERR_INVALID: "AXR_ERR_INVALID", AXR_ERR_OOM: "AXR_ERR_OOM",
    AXR_ERR_HIP: "AXR_ERR_HIP", AXR_ERR_OVERFLOW: "AXR_ERR_OVERFLOW", AXR_ERR_BUDGET: "AXR_ERR_BUDGET",
    AXR_ERR_NO_DEVICE: "AXR_ERR_NO_DEVICE", AXR_ERR_NOT_READY: "AXR_ERR_NOT_READY",
}

AXR_DRAW_ASYNC = 1


class NativeLibraryMissing(RuntimeError):
    """libaxiomr_amd.so has not been built (run __graft_entry__.build())."""


class AxrError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


class AxrMaterial(C.Structure):
    _fields_ = [("diffuse", C.c_void_p), ("metallic", C.c_void_p), ("bump", C.c_void_p),
                ("roughness", C.c_void_p), ("ao", C.c_void_p), ("specular_exponent", C.c_float)]


class AxrGroup(C.Structure):
    _fields_ = [("material", C.c_uint32), ("first_face", C.c_uint32), ("face_count", C.c_uint32)]


class AxrDrawParams(C.Structure):
    _fields_ = [("model", C.c_float * 16), ("view_proj", C.c_float * 16), ("viewport", C.c_float * 16),
                ("camera_pos", C.c_float * 3), ("shader", C.c_int32), ("light_dir", C.c_float * 3),
                ("light_color", C.c_float * 3), ("num_lights", C.c_int32), ("light_pos", (C.c_float * 3) * 4),
                ("light_colors", (C.c_float * 3) * 4), ("flags", C.c_uint32)]


class AxrConfig(C.Structure):
    _fields_ = [("tile_size", C.c_int32), ("raster_threads", C.c_int32), ("bin_capacity", C.c_int32),
                ("small_footprint", C.c_int32), ("raster_mode", C.c_int32), ("reserved", C.c_int32 * 3)]


class AxrKernelStats(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("frame_ms_total", C.c_double), ("stage_ms_total", C.c_double * 6),
                ("passes_skipped", C.c_uint64), ("skip_faults", C.c_uint64)]

    STAGES = ("setup", "clip", "scan", "scatter", "raster", "vertex pass")


class AxrServerConfig(C.Structure):
    _fields_ = [("lanes", C.c_int32), ("max_batch_views", C.c_uint32), ("max_batch_faces", C.c_uint64),
                ("max_batch_tiles", C.c_uint64), ("hbm_budget", C.c_uint64), ("clear_bgra", C.c_uint32),
                ("with_ids", C.c_int32), ("max_batch_pixels", C.c_uint64)]


class AxrServeRequest(C.Structure):
    _fields_ = [("mesh", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32), ("params", AxrDrawParams)]


class AxrServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("submitted", "completed", "failed", "rejected", "batches",
                                           "batched_views", "max_views_in_batch", "launches_overlapped")] + [
        ("max_lanes_busy", C.c_uint32), ("reserved0", C.c_uint32)] + [
        (n, C.c_uint64) for n in ("deferred", "fb_reserved_bytes", "fb_peak_bytes", "fb_in_use_bytes",
                                  "fixed_bytes", "queue_peak", "overflow_retries")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "reserved0"}


class AxrObjGroup(C.Structure):
    _fields_ = [("material", C.c_char * 128), ("first_face", C.c_uint64), ("face_count", C.c_uint64),
                ("material_index", C.c_uint32)]


class AxrObjMaterial(C.Structure):
    _fields_ = [("name", C.c_char * 128), ("ambient", C.c_float * 3), ("diffuse", C.c_float * 3),
                ("specular", C.c_float * 3), ("specular_exponent", C.c_float), ("refractive_index", C.c_float),
                ("dissolve", C.c_float), ("illumination_model", C.c_int32), ("diffuse_map", C.c_char * 512),
                ("metallic_map", C.c_char * 512), ("bump_map", C.c_char * 512), ("roughness_map", C.c_char * 512),
                ("ao_map", C.c_char * 512)]


class AxrCameraInput(C.Structure):
    _fields_ = [("mouse_x", C.c_float), ("mouse_y", C.c_float), ("right_button", C.c_int32), ("keys", C.c_uint32)]


class AxrCamera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("orientation", C.c_float * 4), ("forward", C.c_float * 3),
                ("right", C.c_float * 3), ("up", C.c_float * 3), ("delta_yaw", C.c_float),
                ("delta_pitch", C.c_float), ("velocity", C.c_float * 3), ("fov", C.c_float), ("aspect", C.c_float),
                ("znear", C.c_float), ("zfar", C.c_float), ("mouse_sensitivity", C.c_float),
                ("last_mouse", C.c_float * 2), ("viewport", C.c_int32 * 4), ("view", C.c_float * 16),
                ("proj", C.c_float * 16), ("view_proj", C.c_float * 16), ("view_dirty", C.c_int32),
                ("proj_dirty", C.c_int32)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 1, 2, 3, 4
KEY_W, KEY_S, KEY_A, KEY_D = 1, 2, 4, 8
RASTER_BINNED, RASTER_VISIBILITY = 0, 1

_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
_F = C.POINTER(C.c_float)
_U32 = C.POINTER(C.c_uint32)
_U8 = C.POINTER(C.c_uint8)

# name -> (restype, argtypes)
SIGNATURES = {
    "axr_version": (C.c_char_p, []),
    "axr_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "axr_ctx_create": (C.c_int, [C.c_int, C.c_uint64, _PP]),
    "axr_ctx_destroy": (C.c_int, [_P]),
    "axr_last_error": (C.c_char_p, [_P]),
    "axr_ctx_set_config": (C.c_int, [_P, C.POINTER(AxrConfig)]),
    "axr_ctx_get_config": (C.c_int, [_P, C.POINTER(AxrConfig)]),
    "axr_ctx_set_profiling": (C.c_int, [_P, C.c_int]),
    "axr_ctx_get_stats": (C.c_int, [_P, C.POINTER(AxrKernelStats)]),
    "axr_ctx_memory": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "axr_sync": (C.c_int, [_P]),
    "axr_texture_create": (C.c_int, [_P, _U8, C.c_int, C.c_int, C.c_int, _PP]),
    "axr_texture_destroy": (C.c_int, [_P]),
    "axr_mesh_create": (C.c_int, [_P, _F, C.c_uint64, _U32, C.c_uint64, C.POINTER(AxrGroup), C.c_uint32,
                                  C.POINTER(AxrMaterial), C.c_uint32, _U32, _PP]),
    "axr_mesh_destroy": (C.c_int, [_P]),
    "axr_fb_create": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _PP]),
    "axr_fb_destroy": (C.c_int, [_P]),
    "axr_fb_clear": (C.c_int, [_P, C.c_uint32, C.c_float]),
    "axr_fb_read": (C.c_int, [_P, _U8, _F, _U32]),
    "axr_draw": (C.c_int, [_P, _P, _P, C.POINTER(AxrDrawParams)]),
    "axr_check_view": (C.c_int, [C.c_uint64, C.c_int, C.c_int]),
    "axr_draw_noalloc": (C.c_int, [_P, _P, _P, C.POINTER(AxrDrawParams)]),
    "axr_batch_create": (C.c_int, [_P, C.c_uint32, _PP, _PP, C.POINTER(AxrDrawParams), _PP]),
    "axr_batch_set_clear": (C.c_int, [_P, C.c_int, C.c_uint32, C.c_float]),
    "axr_batch_launch": (C.c_int, [_P]),
    "axr_batch_destroy": (C.c_int, [_P]),
    "axr_server_create": (C.c_int, [_P, C.POINTER(AxrServerConfig), _PP]),
    "axr_server_destroy": (C.c_int, [_P]),
    "axr_server_last_error": (C.c_char_p, [_P]),
    "axr_server_submit": (C.c_int, [_P, C.POINTER(AxrServeRequest), C.POINTER(C.c_uint64)]),
    "axr_server_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_int)]),
    "axr_server_wait": (C.c_int, [_P, C.c_uint64, C.c_double]),
    "axr_server_read": (C.c_int, [_P, C.c_uint64, _U8, _F, _U32]),
    "axr_server_release": (C.c_int, [_P, C.c_uint64]),
    "axr_server_pause": (C.c_int, [_P, C.c_int]),
    "axr_server_get_stats": (C.c_int, [_P, C.POINTER(AxrServerStats)]),
    "axr_server_replay": (C.c_int, [_P, C.c_uint32, C.POINTER(AxrServeRequest), C.POINTER(C.c_double),
                                    C.POINTER(C.c_double)]),
    "axr_obj_load": (C.c_int, [C.c_char_p, _PP]),
    "axr_obj_destroy": (C.c_int, [_P]),
    "axr_obj_last_error": (C.c_char_p, []),
    "axr_obj_counts": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint32)]),
    "axr_obj_arrays": (C.c_int, [_P, _F, _U32]),
    "axr_obj_get_group": (C.c_int, [_P, C.c_uint32, C.POINTER(AxrObjGroup)]),
    "axr_obj_get_material": (C.c_int, [_P, C.c_uint32, C.POINTER(AxrObjMaterial)]),
    "axr_image_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(_U8)]),
    "axr_image_free": (None, [_U8]),
    "axr_image_last_error": (C.c_char_p, []),
    "axr_image_write_png": (C.c_int, [C.c_char_p, C.c_int, C.c_int, _U8, C.c_int]),
    "axr_texture_load": (C.c_int, [_P, C.c_char_p, C.c_int, _PP]),
    "axr_fb_write_png": (C.c_int, [_P, C.c_char_p]),
    "axr_gen_uv_sphere": (C.c_int, [C.c_int, C.c_int, C.c_float, _F, _U32]),
    "axr_gen_icosphere": (C.c_int, [C.c_int, C.c_float, C.c_uint64, _F, _U32]),
    "axr_gen_soup": (C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                               C.c_uint64, _F, _U32, _U32]),
    "axr_gen_checker": (C.c_int, [C.c_int, C.c_int, C.c_int, _U8]),
    "axr_gen_noise_texture": (C.c_int, [C.c_int, C.c_int, C.c_uint64, _U8]),
    "axr_gen_normal_map": (C.c_int, [C.c_int, C.c_int, C.c_float, C.c_uint64, _U8]),
    "axr_camera_matrices": (C.c_int, [_F, _F, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                                      _F, _F, _F, _F]),
    "axr_camera_init": (C.c_int, [C.POINTER(AxrCamera), _F, _F, C.c_float, C.c_float]),
    "axr_camera_handle_input": (C.c_int, [C.POINTER(AxrCamera), C.POINTER(AxrCameraInput)]),
    "axr_camera_update": (C.c_int, [C.POINTER(AxrCamera), C.c_float]),
    "axr_camera_set_position": (C.c_int, [C.POINTER(AxrCamera), _F]),
    "axr_camera_set_target": (C.c_int, [C.POINTER(AxrCamera), _F]),
    "axr_camera_set_fov": (C.c_int, [C.POINTER(AxrCamera), C.c_float]),
    "axr_camera_set_aspect_ratio": (C.c_int, [C.POINTER(AxrCamera), C.c_float]),
    "axr_camera_set_viewport": (C.c_int, [C.POINTER(AxrCamera), C.c_int, C.c_int, C.c_int, C.c_int]),
    "axr_camera_viewport_matrix": (C.c_int, [C.POINTER(AxrCamera), _F]),
    "axr_camera_path": (C.c_int, [C.POINTER(AxrCamera), C.POINTER(AxrCameraInput), _F, C.c_uint64, _F, _F]),
}

_lib = None


def lib() -> C.CDLL:
    """Load the native library once; raise NativeLibraryMissing if it is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NativeLibraryMissing(f"{LIB_PATH} not found; build it with __graft_entry__.build()")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(code: int, ctx=None) -> None:
    if code != AXR_OK:
        msg = ""
        if ctx is not None:
            raw = lib().axr_last_error(ctx)
            msg = raw.decode() if raw else ""
        raise AxrError(code, msg)


def fptr(a):
    return a.ctypes.data_as(_F)


def u32ptr(a):
    return a.ctypes.data_as(_U32)


def u8ptr(a):
    return a.ctypes.data_as(_U8)
