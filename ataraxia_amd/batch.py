"""Batch: many scenes rendered by the persistent wavefront kernel (one launch
per frame pass, all passes replayed from a HIP graph)."""
from __future__ import annotations

import ctypes as C
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import check, lib, ptr
from .scene import FlatScene


def _stats(s: _lib.RenderStats) -> dict:
    return s.as_dict()


class Batch:
    def __init__(self):
        h = C.c_void_p()
        check(lib.atx_batch_create(C.byref(h)))
        self._h = h
        self.scenes: List[FlatScene] = []

    def close(self) -> None:
        if self._h:
            lib.atx_batch_destroy(self._h)
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

    def add_scene(self, scene: FlatScene, spp: Optional[int] = None) -> int:
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
        sid = C.c_uint32()
        check(lib.atx_batch_add_scene(self._h, C.byref(d), C.byref(sid)))
        self.scenes.append(scene)
        return sid.value

    def set_accumulator(self, kind: int) -> None:
        check(lib.atx_batch_set_accumulator(self._h, kind))

    def set_accel(self, accel: int) -> None:
        """ATX_ACCEL_BVH (default) or ATX_ACCEL_NONE (flat loop); same results."""
        check(lib.atx_batch_set_accel(self._h, accel))

    def set_tuning(self, chunk_px: int = 0, workgroups: int = 0, use_graph: bool = True,
                   frames_per_launch: int = 0, waves_per_simd: int = 0) -> None:
        t = _lib.BatchTuning(chunk_px, workgroups, 1 if use_graph else 0, frames_per_launch, waves_per_simd)
        check(lib.atx_batch_set_tuning(self._h, C.byref(t)))

    @staticmethod
    def _tuning(t: _lib.BatchTuning) -> dict:
        return {"chunk_px": t.chunk_px, "workgroups": t.workgroups, "use_graph": bool(t.use_graph),
                "frames_per_launch": t.frames_per_launch, "waves_per_simd": t.waves_per_simd}

    def get_tuning(self) -> dict:
        t = _lib.BatchTuning()
        check(lib.atx_batch_get_tuning(self._h, C.byref(t)))
        return self._tuning(t)

    def autotune(self, cache_path: Optional[str] = None) -> dict:
        t = _lib.BatchTuning()
        check(lib.atx_batch_autotune(self._h, cache_path.encode() if cache_path else None, C.byref(t)))
        return self._tuning(t)

    def render(self, max_frames: Optional[int] = None) -> dict:
        if max_frames is None:
            check(lib.atx_batch_render(self._h))
        else:
            check(lib.atx_batch_render_frames(self._h, max_frames))
        return self.stats()

    def capture(self, scene_id: int, max_bounces: int) -> None:
        check(lib.atx_batch_capture(self._h, scene_id, max_bounces))

    def get_capture(self, scene_id: int, max_bounces: int):
        sc = self.scenes[scene_id]
        n = sc.camera.width * sc.camera.height
        rad = np.zeros((n, 4), np.float32)
        hit = np.zeros((n, max_bounces), np.int32)
        sh = np.zeros((n, max_bounces), np.int32)
        check(lib.atx_batch_get_capture(self._h, scene_id, ptr(rad), ptr(hit), ptr(sh)))
        return rad, hit, sh

    def get_image(self, scene_id: int) -> np.ndarray:
        sc = self.scenes[scene_id]
        img = np.zeros(sc.camera.width * sc.camera.height, np.uint32)
        check(lib.atx_batch_get_image(self._h, scene_id, ptr(img), img.size))
        return img.reshape(sc.camera.height, sc.camera.width)

    def get_accumulation(self, scene_id: int) -> np.ndarray:
        sc = self.scenes[scene_id]
        acc = np.zeros(sc.camera.width * sc.camera.height * 4, np.float32)
        check(lib.atx_batch_get_accumulation(self._h, scene_id, ptr(acc), acc.size))
        return acc.reshape(sc.camera.height, sc.camera.width, 4)

    def stats(self) -> dict:
        s = _lib.RenderStats()
        check(lib.atx_batch_get_stats(self._h, C.byref(s)))
        return _stats(s)
