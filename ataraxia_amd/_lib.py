"""ctypes binding of libatx.so (the C ABI declared in include/atx.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C ataraxia_amd/csrc``).  Nothing here falls back to a CPU path: if
the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libatx.so")


class AtxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"atx error {code}: {msg}")
        self.code = code


ATX_OK = 0
ATX_ERR_INVALID = -1
ATX_ERR_HIP = -2
ATX_ERR_NOMEM = -3
ATX_ERR_STATE = -4
ATX_ERR_IO = -5
ATX_ERR_BUDGET = -6
ATX_ERR_NOTFOUND = -7
ATX_ERR_NODEVICE = -8

ATX_ACCUM_NAIVE = 0
ATX_ACCUM_KAHAN = 1
ATX_KERNEL_PLAIN = 0
ATX_KERNEL_WAVEFRONT = 1
ATX_ACCEL_BVH = 0
ATX_ACCEL_NONE = 1
ATX_REQ_QUEUED, ATX_REQ_RUNNING, ATX_REQ_DONE, ATX_REQ_FAILED = 0, 1, 2, 3


class Sphere(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("radius", C.c_float), ("material", C.c_int32)]


class Material(C.Structure):
    _fields_ = [("albedo", C.c_float * 3), ("roughness", C.c_float), ("metallic", C.c_float),
                ("F0", C.c_float * 3), ("emission_color", C.c_float * 3),
                ("emission_intensity", C.c_float), ("id", C.c_int32)]


class Light(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("color", C.c_float * 3), ("intensity", C.c_float)]


class Settings(C.Structure):
    _fields_ = [("accumulation", C.c_uint8), ("sky_light", C.c_uint8), ("reserved", C.c_uint8 * 2),
                ("max_bounces", C.c_int32)]


class CameraDesc(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("fov_deg", C.c_float),
                ("near_clip", C.c_float), ("far_clip", C.c_float), ("width", C.c_uint32),
                ("height", C.c_uint32)]


class SceneDesc(C.Structure):
    _fields_ = [("spheres", C.c_void_p), ("num_spheres", C.c_uint64), ("materials", C.c_void_p),
                ("num_materials", C.c_uint64), ("lights", C.c_void_p), ("num_lights", C.c_uint64),
                ("settings", Settings), ("camera", CameraDesc), ("spp", C.c_uint32), ("reserved", C.c_uint32)]


class BatchTuning(C.Structure):
    _fields_ = [("chunk_px", C.c_uint32), ("workgroups", C.c_uint32), ("use_graph", C.c_uint32),
                ("frames_per_launch", C.c_uint32), ("waves_per_simd", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("frames_per_tick", C.c_uint32), ("chunk_px", C.c_uint32),
                ("max_active", C.c_uint32), ("kahan", C.c_uint32), ("accel", C.c_uint32),
                ("tick_mpaths", C.c_uint32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("admitted", C.c_uint64), ("completed", C.c_uint64),
                ("rejected", C.c_uint64), ("failed", C.c_uint64), ("queued", C.c_uint64), ("active", C.c_uint64),
                ("samples", C.c_uint64), ("ticks", C.c_uint64), ("hbm_in_use", C.c_uint64),
                ("hbm_high_water", C.c_uint64), ("device_used_high_water", C.c_uint64),
                ("gpu_busy_ms", C.c_double), ("admit_ms", C.c_double), ("retire_ms", C.c_double),
                ("wait_ms", C.c_double), ("release_ms", C.c_double), ("gap_ms", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("state", C.c_int32), ("frames_done", C.c_uint32), ("t_submit_s", C.c_double),
                ("t_admit_s", C.c_double), ("t_done_s", C.c_double)]


class RenderStats(C.Structure):
    _fields_ = [("samples", C.c_uint64), ("rays", C.c_uint64), ("sphere_tests", C.c_uint64),
                ("exact_retraces", C.c_uint64), ("box_tests", C.c_uint64), ("gpu_ms", C.c_double),
                ("launches", C.c_uint64), ("path_ms", C.c_double), ("reused_rays", C.c_uint64)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


assert C.sizeof(Sphere) == 20 and C.sizeof(Material) == 52 and C.sizeof(Light) == 28
assert C.sizeof(Settings) == 8

# numpy views of the same layouts (packed, little endian)
SPHERE_DTYPE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("material", "<i4")])
MATERIAL_DTYPE = np.dtype([("albedo", "<f4", 3), ("roughness", "<f4"), ("metallic", "<f4"),
                           ("F0", "<f4", 3), ("emission_color", "<f4", 3),
                           ("emission_intensity", "<f4"), ("id", "<i4")])
LIGHT_DTYPE = np.dtype([("position", "<f4", 3), ("color", "<f4", 3), ("intensity", "<f4")])
assert SPHERE_DTYPE.itemsize == 20 and MATERIAL_DTYPE.itemsize == 52 and LIGHT_DTYPE.itemsize == 28

_P = C.c_void_p
_U32 = C.c_uint32
_U64 = C.c_uint64
_I = C.c_int

# name -> (restype, argtypes); every name here is declared in include/atx.h
PROTOTYPES = {
    "atx_last_error": (C.c_char_p, []),
    "atx_version": (C.c_char_p, []),
    "atx_device_count": (_I, [C.POINTER(_I)]),
    "atx_device_synchronize": (_I, []),
    "atx_renderer_create": (_I, [C.POINTER(_P)]),
    "atx_renderer_destroy": (_I, [_P]),
    "atx_renderer_on_resize": (_I, [_P, _U32, _U32]),
    "atx_renderer_set_scene": (_I, [_P, _P, _U64, _P, _U64, _P, _U64]),
    "atx_renderer_set_settings": (_I, [_P, C.POINTER(Settings)]),
    "atx_renderer_get_settings": (_I, [_P, C.POINTER(Settings)]),
    "atx_renderer_reset_frame_index": (_I, [_P]),
    "atx_renderer_get_frame_index": (_I, [_P, C.POINTER(_U32)]),
    "atx_renderer_set_accumulator": (_I, [_P, _I]),
    "atx_renderer_render": (_I, [_P, C.POINTER(CameraDesc)]),
    "atx_renderer_render_frames": (_I, [_P, C.POINTER(CameraDesc), _U32]),
    "atx_renderer_get_image": (_I, [_P, _P, _U64]),
    "atx_renderer_get_accumulation": (_I, [_P, _P, _U64]),
    "atx_renderer_get_stats": (_I, [_P, C.POINTER(RenderStats)]),
    "atx_renderer_render_debug": (_I, [_P, C.POINTER(CameraDesc), _P, _P, _P, _U32]),
    "atx_renderer_set_kernel": (_I, [_P, _I]),
    "atx_renderer_set_accel": (_I, [_P, _I]),
    "atx_renderer_reserve": (_I, [_P, _U32]),
    "atx_render_frames_noalloc": (_I, [_P, _U64, _P, _U64, _P, _U64, C.POINTER(Settings), C.POINTER(CameraDesc),
                                       _U32, _U32, _I, _I, _P, _P, _P, _P]),
    "atx_batch_create": (_I, [C.POINTER(_P)]),
    "atx_batch_destroy": (_I, [_P]),
    "atx_batch_add_scene": (_I, [_P, C.POINTER(SceneDesc), C.POINTER(_U32)]),
    "atx_batch_set_accumulator": (_I, [_P, _I]),
    "atx_batch_set_accel": (_I, [_P, _I]),
    "atx_batch_set_tuning": (_I, [_P, C.POINTER(BatchTuning)]),
    "atx_batch_get_tuning": (_I, [_P, C.POINTER(BatchTuning)]),
    "atx_batch_autotune": (_I, [_P, C.c_char_p, C.POINTER(BatchTuning)]),
    "atx_batch_render": (_I, [_P]),
    "atx_batch_render_frames": (_I, [_P, _U32]),
    "atx_batch_capture": (_I, [_P, _U32, _U32]),
    "atx_batch_get_capture": (_I, [_P, _U32, _P, _P, _P]),
    "atx_batch_get_image": (_I, [_P, _U32, _P, _U64]),
    "atx_batch_get_accumulation": (_I, [_P, _U32, _P, _U64]),
    "atx_batch_get_stats": (_I, [_P, C.POINTER(RenderStats)]),
    "atx_server_create": (_I, [C.POINTER(_P), C.POINTER(ServerConfig)]),
    "atx_server_destroy": (_I, [_P]),
    "atx_server_submit": (_I, [_P, C.POINTER(SceneDesc), C.POINTER(_U64)]),
    "atx_server_get_info": (_I, [_P, _U64, C.POINTER(RequestInfo)]),
    "atx_server_wait": (_I, [_P, _U64, C.c_double]),
    "atx_server_get_result": (_I, [_P, _U64, _P, _U64, _P, _U64, _I]),
    "atx_server_get_stats": (_I, [_P, C.POINTER(ServerStats)]),
    "atx_scene_create": (_I, [C.POINTER(_P)]),
    "atx_scene_destroy": (_I, [_P]),
    "atx_scene_load_json": (_I, [C.c_char_p, C.POINTER(_P)]),
    "atx_scene_parse_json": (_I, [C.c_char_p, _U64, C.POINTER(_P)]),
    "atx_scene_save_json": (_I, [_P, C.c_char_p]),
    "atx_scene_to_json": (_I, [_P, C.c_char_p, _U64, C.POINTER(_U64)]),
    "atx_scene_add_node": (_I, [_P, _U32, C.c_char_p, _P, _P, _P, C.POINTER(_U32)]),
    "atx_scene_set_node_transform": (_I, [_P, _U32, _P, _P, _P]),
    "atx_scene_add_sphere": (_I, [_P, _U32, C.POINTER(Sphere)]),
    "atx_scene_add_material": (_I, [_P, C.POINTER(Material)]),
    "atx_scene_add_light": (_I, [_P, C.POINTER(Light)]),
    "atx_scene_set_settings": (_I, [_P, C.POINTER(Settings)]),
    "atx_scene_get_settings": (_I, [_P, C.POINTER(Settings)]),
    "atx_scene_set_camera": (_I, [_P, _P, _P, C.c_float]),
    "atx_scene_get_camera": (_I, [_P, _P, _P, C.POINTER(C.c_float)]),
    "atx_scene_counts": (_I, [_P, C.POINTER(_U64), C.POINTER(_U64), C.POINTER(_U64), C.POINTER(_U64)]),
    "atx_scene_get_node": (_I, [_P, _U32, C.c_char_p, _U64, _P, _P, _P, C.POINTER(_U32), C.POINTER(_U64)]),
    "atx_scene_get_node_spheres": (_I, [_P, _U32, _P, _U64, C.POINTER(_U64)]),
    "atx_scene_flatten": (_I, [_P, _P, _U64, C.POINTER(_U64)]),
    "atx_scene_get_materials": (_I, [_P, _P, _U64, C.POINTER(_U64)]),
    "atx_scene_get_lights": (_I, [_P, _P, _U64, C.POINTER(_U64)]),
    "atx_write_ppm": (_I, [C.c_char_p, _P, _U32, _U32, _I]),
    "atx_camera_ray_directions": (_I, [C.POINTER(CameraDesc), _P, _U64]),
    "atx_camera_matrices": (_I, [C.POINTER(CameraDesc), _P, _P]),
}


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `make -C ataraxia_amd/csrc`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def last_error() -> str:
    return (lib.atx_last_error() or b"").decode(errors="replace")


def check(rc: int) -> None:
    if rc != ATX_OK:
        raise AtxError(rc, last_error())


def ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def device_count() -> int:
    n = C.c_int(0)
    rc = lib.atx_device_count(C.byref(n))
    return n.value if rc == ATX_OK else 0
