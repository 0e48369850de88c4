"""ctypes view of the CPU oracle (oracle/libatx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() to
compare its result, and by bench.py's cpu_baseline leg.  The product package
(ataraxia_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libatx_oracle.so")

_P = C.c_void_p
_F = C.c_float
_U32 = C.c_uint32
_U64 = C.c_uint64


class OraScene(C.Structure):
    _fields_ = [("spheres", _P), ("num_spheres", _U64), ("materials", _P), ("num_materials", _U64),
                ("lights", _P), ("num_lights", _U64)]


class OraSettings(C.Structure):
    _fields_ = [("accumulation", C.c_uint8), ("sky_light", C.c_uint8), ("pad", C.c_uint8 * 2),
                ("max_bounces", C.c_int32)]


class OraCamera(C.Structure):
    _fields_ = [("position", _F * 3), ("direction", _F * 3), ("fov_deg", _F), ("near_clip", _F),
                ("far_clip", _F), ("width", _U32), ("height", _U32)]


class OraPathDebug(C.Structure):
    _fields_ = [("hit_ids", _P), ("shadow_ids", _P), ("rays", _U64), ("bounces", C.c_int32),
                ("seeds", _P), ("light_ids", _P), ("rr", _P), ("dirs", _P)]


_PROTOS = {
    "ora_pcg_hash": (_U32, [_U32]),
    "ora_pcg_float": (_F, [C.POINTER(_U32)]),
    "ora_sin": (_F, [_F]),
    "ora_cos": (_F, [_F]),
    "ora_pow5": (_F, [_F]),
    "ora_perspective": (None, [_F, _F, _F, _F, _P]),
    "ora_look_at": (None, [_P, _P, _P, _P]),
    "ora_inverse": (None, [_P, _P]),
    "ora_camera_matrices": (None, [C.POINTER(OraCamera), _P, _P]),
    "ora_ray_direction": (None, [_P, _P, _U32, _U32, _U32, _U32, _P]),
    "ora_ray_directions": (None, [C.POINTER(OraCamera), _P]),
    "ora_trace_ray": (C.c_int32, [_P, _P, _P, _U64, C.POINTER(_F)]),
    "ora_cook_torrance": (None, [_P, _P, _F, _F, _P, _P, _P, _P]),
    "ora_sample_cosine": (None, [_P, C.POINTER(_U32), _P]),
    "ora_sample_ggx": (None, [_P, _F, C.POINTER(_U32), _P]),
    "ora_per_pixel": (None, [_U32, _U32, _U32, _P, _P, C.POINTER(OraScene), _U32,
                             C.POINTER(OraSettings), _P, C.POINTER(OraPathDebug)]),
    "ora_render_frame": (None, [C.POINTER(OraCamera), C.POINTER(OraScene), C.POINTER(OraSettings), _U32,
                                _P, _P, _P, _P, _U32, _U32, C.POINTER(_U64)]),
    "ora_render_frames_mt": (C.c_int, [C.POINTER(OraCamera), C.POINTER(OraScene), C.POINTER(OraSettings),
                                       _U32, _U32, _P, _P, _P, C.c_int, C.POINTER(_U64)]),
    "ora_rgba8": (_U32, [_P]),
    "ora_render_frame_debug": (None, [C.POINTER(OraCamera), C.POINTER(OraScene), C.POINTER(OraSettings), _U32,
                                      _P, _P, _P, _U32, _P, _U32, _U32]),
    "ora_accumulate_pixels": (C.c_int, [C.POINTER(OraCamera), C.POINTER(OraScene), C.POINTER(OraSettings), _U32,
                                        _U32, _P, _U64, C.c_int, _P, _P, C.c_int]),
    "ora_accumulate_pixels_f64": (C.c_int, [C.POINTER(OraCamera), C.POINTER(OraScene), C.POINTER(OraSettings),
                                            _U32, _U32, _P, _U64, _P, C.c_int]),
    "ora_node_local_transform": (None, [_P, _P, _P, _P]),
    "ora_mat4_mul": (None, [_P, _P, _P]),
    "ora_flatten_sphere": (None, [_P, _P, _P]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data if a.size else None


class Scene:
    """Keeps the numpy arrays alive while the oracle reads them."""

    def __init__(self, spheres, materials, lights):
        self.spheres = np.ascontiguousarray(spheres)
        self.materials = np.ascontiguousarray(materials)
        self.lights = np.ascontiguousarray(lights)
        self.c = OraScene(_p(self.spheres), len(self.spheres), _p(self.materials), len(self.materials),
                          _p(self.lights), len(self.lights))


def settings(max_bounces: int, sky_light: bool = False, accumulation: bool = True) -> OraSettings:
    s = OraSettings()
    s.accumulation = 1 if accumulation else 0
    s.sky_light = 1 if sky_light else 0
    s.max_bounces = max_bounces
    return s


def camera(position, direction, width, height, fov=45.0, near=0.1, far=100.0) -> OraCamera:
    c = OraCamera()
    c.position[:] = [float(np.float32(v)) for v in position]
    c.direction[:] = [float(np.float32(v)) for v in direction]
    c.fov_deg, c.near_clip, c.far_clip = fov, near, far
    c.width, c.height = width, height
    return c


def camera_from(cam) -> OraCamera:
    """From an ataraxia_amd.scene.Camera-like object (duck typed)."""
    return camera(cam.position, cam.direction, cam.width, cam.height, cam.fov, cam.near_clip, cam.far_clip)


def ray_directions(cam: OraCamera) -> np.ndarray:
    out = np.zeros(cam.width * cam.height * 3, dtype=np.float32)
    lib.ora_ray_directions(C.byref(cam), _p(out))
    return out.reshape(cam.height, cam.width, 3)


def camera_matrices(cam: OraCamera):
    ip = np.zeros(16, np.float32)
    iv = np.zeros(16, np.float32)
    lib.ora_camera_matrices(C.byref(cam), _p(ip), _p(iv))
    return ip, iv


def per_pixel(x, y, width, direction, cam_pos, scene: Scene, frame, sett: OraSettings):
    """Returns (rgba float32[4], hit_ids, shadow_ids, rays)."""
    nb = max(1, sett.max_bounces)
    hit = np.full(nb, -2, np.int32)
    sh = np.full(nb, -2, np.int32)
    dbg = OraPathDebug(_p(hit), _p(sh), 0, 0)
    out = np.zeros(4, np.float32)
    d = np.ascontiguousarray(direction, dtype=np.float32)
    o = np.ascontiguousarray(cam_pos, dtype=np.float32)
    lib.ora_per_pixel(x, y, width, _p(d), _p(o), C.byref(scene.c), frame, C.byref(sett), _p(out), C.byref(dbg))
    return out, hit, sh, dbg.rays


def per_pixel_trace(x, y, width, direction, cam_pos, scene: Scene, frame, sett: OraSettings) -> dict:
    """One path with its per-bounce log: hit and shadow ids, the seed after
    `seed += i`, the light index, roulette p and u, and the traced direction."""
    nb = max(1, sett.max_bounces)
    log = {"hit": np.full(nb, -2, np.int32), "shadow": np.full(nb, -2, np.int32),
           "seed": np.zeros(nb, np.uint32), "light": np.full(nb, -1, np.int32),
           "rr": np.full((nb, 2), np.nan, np.float32), "dir": np.full((nb, 3), np.nan, np.float32)}
    dbg = OraPathDebug(_p(log["hit"]), _p(log["shadow"]), 0, 0, _p(log["seed"]), _p(log["light"]),
                       _p(log["rr"]), _p(log["dir"]))
    out = np.zeros(4, np.float32)
    d = np.ascontiguousarray(direction, dtype=np.float32)
    o = np.ascontiguousarray(cam_pos, dtype=np.float32)
    lib.ora_per_pixel(x, y, width, _p(d), _p(o), C.byref(scene.c), frame, C.byref(sett), _p(out), C.byref(dbg))
    log["rgba"] = out
    log["rays"] = dbg.rays
    log["bounces"] = dbg.bounces
    return log


def render_frame(cam: OraCamera, scene: Scene, sett: OraSettings, frame: int, accum=None, comp=None,
                 want_radiance=True, rows=None):
    """One frame over `rows` (default all). Returns (accum, comp, image, radiance, rays)."""
    n = cam.width * cam.height
    if accum is None:
        accum = np.zeros((n, 4), np.float32)
    img = np.zeros(n, np.uint32)
    rad = np.zeros((n, 4), np.float32) if want_radiance else None
    rays = _U64(0)
    r0, r1 = rows if rows is not None else (0, cam.height)
    lib.ora_render_frame(C.byref(cam), C.byref(scene.c), C.byref(sett), frame, _p(accum),
                         _p(comp) if comp is not None else None, _p(img),
                         _p(rad) if rad is not None else None, r0, r1, C.byref(rays))
    return accum, comp, img, rad, rays.value


def render_frames_mt(cam: OraCamera, scene: Scene, sett: OraSettings, first_frame: int, frames: int,
                     threads: int, accum=None, comp=None, kahan=False):
    n = cam.width * cam.height
    if accum is None:
        accum = np.zeros((n, 4), np.float32)
    if kahan and comp is None:
        comp = np.zeros((n, 4), np.float32)
    img = np.zeros(n, np.uint32)
    rays = _U64(0)
    rc = lib.ora_render_frames_mt(C.byref(cam), C.byref(scene.c), C.byref(sett), first_frame, frames,
                                  _p(accum), _p(comp) if comp is not None else None, _p(img), threads,
                                  C.byref(rays))
    if rc != 0:
        raise RuntimeError("oracle thread creation failed")
    return accum, comp, img, rays.value


def render_frame_debug(cam: OraCamera, scene: Scene, sett: OraSettings, frame: int, max_bounces: int,
                       rows=None):
    """Per-pixel radiance, hit ids, shadow ids and ray counts of one frame."""
    n = cam.width * cam.height
    rad = np.zeros((n, 4), np.float32)
    hit = np.full((n, max_bounces), -2, np.int32)
    sh = np.full((n, max_bounces), -2, np.int32)
    rays = np.zeros(n, np.uint32)
    r0, r1 = rows if rows is not None else (0, cam.height)
    lib.ora_render_frame_debug(C.byref(cam), C.byref(scene.c), C.byref(sett), frame, _p(rad), _p(hit), _p(sh),
                               max_bounces, _p(rays), r0, r1)
    return rad, hit, sh, rays


def accumulate_pixels(cam: OraCamera, scene: Scene, sett: OraSettings, first_frame: int, frames: int, pixels,
                      kahan: bool = False, threads: int = 8):
    """Accumulation (4 floats per listed pixel) and RGBA8 after `frames`
    frames from first_frame, for pixel indices x + y*width."""
    px = np.ascontiguousarray(pixels, dtype=np.uint32)
    acc = np.zeros((len(px), 4), np.float32)
    img = np.zeros(len(px), np.uint32)
    if len(px):
        rc = lib.ora_accumulate_pixels(C.byref(cam), C.byref(scene.c), C.byref(sett), first_frame, frames,
                                       _p(px), len(px), 1 if kahan else 0, _p(acc), _p(img), threads)
        if rc != 0:
            raise RuntimeError("oracle thread creation failed")
    return acc, img


def accumulate_pixels_f64(cam: OraCamera, scene: Scene, sett: OraSettings, first_frame: int, frames: int, pixels,
                          threads: int = 8):
    """The float64 sum of the same float32 samples (4 doubles per listed pixel)."""
    px = np.ascontiguousarray(pixels, dtype=np.uint32)
    acc = np.zeros((len(px), 4), np.float64)
    if len(px):
        rc = lib.ora_accumulate_pixels_f64(C.byref(cam), C.byref(scene.c), C.byref(sett), first_frame, frames,
                                           _p(px), len(px), _p(acc), threads)
        if rc != 0:
            raise RuntimeError("oracle thread creation failed")
    return acc


def pcg_hash(seed: int) -> int:
    return lib.ora_pcg_hash(seed & 0xFFFFFFFF)
