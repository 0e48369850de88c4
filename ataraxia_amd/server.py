"""Server: continuous-batching render service (C ABI atx_server_*)."""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, lib, ptr
from .scene import FlatScene

REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED = 0, 1, 2, 3


def scene_desc(scene: FlatScene, spp: Optional[int] = None):
    """Returns (desc, keep-alive arrays)."""
    sph = np.ascontiguousarray(scene.spheres)
    mat = np.ascontiguousarray(scene.materials)
    lig = np.ascontiguousarray(scene.lights)
    d = _lib.SceneDesc()
    d.spheres = ptr(sph) if len(sph) else None
    d.num_spheres = len(sph)
    d.materials = ptr(mat) if len(mat) else None
    d.num_materials = len(mat)
    d.lights = ptr(lig) if len(lig) else None
    d.num_lights = len(lig)
    d.settings = scene.settings.to_c()
    d.camera = scene.camera.to_c()
    d.spp = int(spp if spp is not None else scene.spp)
    return d, (sph, mat, lig)


class Server:
    def __init__(self, hbm_budget_bytes: int = 0, frames_per_tick: int = 0, chunk_px: int = 0,
                 max_active: int = 0, kahan: bool = False, accel: int = _lib.ATX_ACCEL_BVH, tick_mpaths: int = 0):
        cfg = _lib.ServerConfig(hbm_budget_bytes, frames_per_tick, chunk_px, max_active, 1 if kahan else 0, accel,
                                tick_mpaths)
        h = C.c_void_p()
        check(lib.atx_server_create(C.byref(h), C.byref(cfg)))
        self._h = h
        self.sizes = {}

    def close(self) -> None:
        if self._h:
            lib.atx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, scene: FlatScene, spp: Optional[int] = None) -> int:
        d, keep = scene_desc(scene, spp)
        rid = C.c_uint64()
        check(lib.atx_server_submit(self._h, C.byref(d), C.byref(rid)))
        del keep
        self.sizes[rid.value] = (scene.camera.width, scene.camera.height)
        return rid.value

    def info(self, rid: int) -> dict:
        i = _lib.RequestInfo()
        check(lib.atx_server_get_info(self._h, rid, C.byref(i)))
        return {"state": i.state, "frames_done": i.frames_done, "t_submit": i.t_submit_s,
                "t_admit": i.t_admit_s, "t_done": i.t_done_s}

    def wait(self, rid: int, timeout_s: float = -1.0) -> None:
        check(lib.atx_server_wait(self._h, rid, timeout_s))

    def result(self, rid: int, release: bool = False):
        w, h = self.sizes[rid]
        img = np.zeros(w * h, np.uint32)
        acc = np.zeros(w * h * 4, np.float32)
        check(lib.atx_server_get_result(self._h, rid, ptr(img), img.size, ptr(acc), acc.size, 1 if release else 0))
        return img.reshape(h, w), acc.reshape(h, w, 4)

    def stats(self) -> dict:
        s = _lib.ServerStats()
        check(lib.atx_server_get_stats(self._h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _lib.ServerStats._fields_}
