"""Renderer: Python view of the C ABI's single-scene renderer.

Method names and semantics follow the reference's Renderer
(Engine/include/Renderer.h:18-31, Engine/src/Renderer.cu:98-249):
``on_resize`` reallocates and resets the frame index, ``render(camera,
scene)`` re-uploads the scene when the scene object changes or the frame
index is 1, renders one frame, and advances the frame index when
accumulation is on.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Tuple

import numpy as np

from . import _lib
from ._lib import check, lib, ptr
from .scene import Camera, FlatScene, Settings


class Renderer:
    def __init__(self):
        h = C.c_void_p()
        check(lib.atx_renderer_create(C.byref(h)))
        self._h = h
        self._scene_id: Optional[int] = None
        self._scene_ref: Optional[FlatScene] = None
        self.width = 0
        self.height = 0

    def close(self) -> None:
        if self._h:
            lib.atx_renderer_destroy(self._h)
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

    # -- Renderer::onResize ---------------------------------------------------
    def on_resize(self, width: int, height: int) -> None:
        check(lib.atx_renderer_on_resize(self._h, width, height))
        self.width, self.height = int(width), int(height)

    # -- scene upload (Renderer::allocateDeviceMemory) -------------------------
    def set_scene(self, scene: FlatScene) -> None:
        sph = np.ascontiguousarray(scene.spheres)
        mat = np.ascontiguousarray(scene.materials)
        lig = np.ascontiguousarray(scene.lights)
        check(lib.atx_renderer_set_scene(self._h, ptr(sph) if len(sph) else None, len(sph),
                                         ptr(mat) if len(mat) else None, len(mat),
                                         ptr(lig) if len(lig) else None, len(lig)))
        self._scene_ref = scene

    # -- settings ----------------------------------------------------------------
    def get_settings(self) -> Settings:
        s = _lib.Settings()
        check(lib.atx_renderer_get_settings(self._h, C.byref(s)))
        return Settings(bool(s.accumulation), bool(s.sky_light), int(s.max_bounces))

    def set_settings(self, settings: Settings) -> None:
        check(lib.atx_renderer_set_settings(self._h, C.byref(settings.to_c())))

    def reset_frame_index(self) -> None:
        check(lib.atx_renderer_reset_frame_index(self._h))

    @property
    def frame_index(self) -> int:
        v = C.c_uint32()
        check(lib.atx_renderer_get_frame_index(self._h, C.byref(v)))
        return v.value

    def set_accumulator(self, kind: int) -> None:
        check(lib.atx_renderer_set_accumulator(self._h, kind))

    def set_kernel(self, kind: int) -> None:
        """ATX_KERNEL_WAVEFRONT (the default, persistent) or ATX_KERNEL_PLAIN (one lane per pixel)."""
        check(lib.atx_renderer_set_kernel(self._h, kind))

    def set_accel(self, accel: int) -> None:
        """ATX_ACCEL_BVH (default) or ATX_ACCEL_NONE (flat loop); same results."""
        check(lib.atx_renderer_set_accel(self._h, accel))

    # -- Renderer::Render ----------------------------------------------------------
    def render(self, camera: Camera, scene: FlatScene) -> None:
        if self._scene_ref is not scene or self.frame_index == 1:
            self.set_scene(scene)
        check(lib.atx_renderer_render(self._h, C.byref(camera.to_c())))

    def reserve(self, frames: int) -> None:
        """Allocates the device buffers that render calls of up to `frames`
        frames at the current size need, so that they allocate nothing."""
        check(lib.atx_renderer_reserve(self._h, frames))

    def render_frames(self, camera: Camera, frames: int) -> None:
        check(lib.atx_renderer_render_frames(self._h, C.byref(camera.to_c()), frames))

    def render_debug(self, camera: Camera, max_bounces: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        n = self.width * self.height
        rad = np.zeros((n, 4), dtype=np.float32)
        hit = np.zeros((n, max_bounces), dtype=np.int32)
        sh = np.zeros((n, max_bounces), dtype=np.int32)
        check(lib.atx_renderer_render_debug(self._h, C.byref(camera.to_c()), ptr(rad), ptr(hit), ptr(sh),
                                            max_bounces))
        return rad, hit, sh

    # -- outputs -------------------------------------------------------------------
    def get_image(self) -> np.ndarray:
        img = np.zeros(self.width * self.height, dtype=np.uint32)
        check(lib.atx_renderer_get_image(self._h, ptr(img), img.size))
        return img.reshape(self.height, self.width)

    def get_accumulation(self) -> np.ndarray:
        acc = np.zeros(self.width * self.height * 4, dtype=np.float32)
        check(lib.atx_renderer_get_accumulation(self._h, ptr(acc), acc.size))
        return acc.reshape(self.height, self.width, 4)

    def stats(self) -> dict:
        """Counters of the last render call: samples, rays (every ray the
        reference traces) and reused_rays (those answered from the
        first-bounce cache), sphere and box tests, exact retraces, launches,
        and HIP-event times of the call (gpu_ms) and of the path kernel
        (path_ms)."""
        s = _lib.RenderStats()
        check(lib.atx_renderer_get_stats(self._h, C.byref(s)))
        return s.as_dict()


def camera_ray_directions(camera: Camera) -> np.ndarray:
    out = np.zeros(camera.width * camera.height * 3, dtype=np.float32)
    check(lib.atx_camera_ray_directions(C.byref(camera.to_c()), ptr(out), out.size))
    return out.reshape(camera.height, camera.width, 3)


def camera_matrices(camera: Camera) -> Tuple[np.ndarray, np.ndarray]:
    ip = np.zeros(16, dtype=np.float32)
    iv = np.zeros(16, dtype=np.float32)
    check(lib.atx_camera_matrices(C.byref(camera.to_c()), ptr(ip), ptr(iv)))
    return ip, iv
