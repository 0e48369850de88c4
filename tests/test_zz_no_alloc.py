"""atx_render_frames_noalloc (the allocation-free entry point) against
atx_renderer_render_frames with the plain kernel, through the C ABI.

The entry point allocates nothing, so the test brings the device buffers
(allocated with the HIP runtime directly) and compares accumulation and
image bit for bit with the renderer's on two calls in a row, the second
with a smaller viewport and scene, and with a continuation call that starts
at a later frame index.
"""
import ctypes as C

import numpy as np
import pytest

from ataraxia_amd import ATX_ACCUM_KAHAN, ATX_ACCUM_NAIVE, ATX_KERNEL_PLAIN, Renderer, Settings
from ataraxia_amd import _lib
from ataraxia_amd.scene import Camera, FlatScene, SplitMix64, random_scene

pytestmark = pytest.mark.gpu

_hip = C.CDLL("libamdhip64.so")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
H2D, D2H = 1, 2


class DeviceArray:
    """A device allocation made by the test (the entry point makes none)."""

    def __init__(self, nbytes: int):
        self.nbytes = nbytes
        self.p = C.c_void_p()
        assert _hip.hipMalloc(C.byref(self.p), nbytes) == 0
        assert _hip.hipMemset(self.p, 0x7F, nbytes) == 0   # garbage: frame 1 must not read it

    def read(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        assert _hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            _hip.hipFree(self.p)
            self.p = None


def scene(seed, n_spheres, n_lights, w, h, bounces, sky=False):
    sph, mat, lig = random_scene(SplitMix64(seed), n_spheres, n_lights, 10.0, (0.0, 4.0), (8.0, 12.0))
    cam = Camera.looking_at((0.0, 5.0, 24.0), width=w, height=h)
    return FlatScene(sph, mat, lig, Settings(True, sky, bounces), cam, name=f"noalloc{seed}")


def noalloc(sc: FlatScene, first_frame: int, frames: int, accel: int, kahan: bool, bufs):
    cam = sc.camera
    sph, mat, lig = (np.ascontiguousarray(a) for a in (sc.spheres, sc.materials, sc.lights))
    print(f"case noalloc {sc.name} {cam.width}x{cam.height} spheres={len(sph)} frames {first_frame}+{frames} "
          f"accel={accel} kahan={kahan}", flush=True)
    acc, comp, img = bufs
    rc = _lib.lib.atx_render_frames_noalloc(
        _lib.ptr(sph), len(sph), _lib.ptr(mat), len(mat), _lib.ptr(lig) if len(lig) else None, len(lig),
        C.byref(sc.settings.to_c()), C.byref(cam.to_c()), first_frame, frames, accel,
        ATX_ACCUM_KAHAN if kahan else ATX_ACCUM_NAIVE, acc.p, comp.p if kahan else None, img.p, None)
    _lib.check(rc)
    n = cam.width * cam.height
    return acc.read(np.float32, (n, 4)), img.read(np.uint32, n)


def reference_entry(sc: FlatScene, frames_list, accel: int, kahan: bool):
    """atx_renderer_render_frames with the plain kernel, frames in chunks."""
    cam = sc.camera
    out = []
    with Renderer() as r:
        r.set_kernel(ATX_KERNEL_PLAIN)
        r.set_accel(accel)
        r.on_resize(cam.width, cam.height)
        r.set_settings(sc.settings)
        if kahan:
            r.set_accumulator(ATX_ACCUM_KAHAN)
        r.set_scene(sc)
        for k in frames_list:
            r.render_frames(cam, k)
            out.append((r.get_accumulation(), r.get_image()))
    return out


@pytest.mark.parametrize("accel", [_lib.ATX_ACCEL_BVH, _lib.ATX_ACCEL_NONE], ids=["bvh", "flat"])
def test_noalloc_equals_renderer_two_calls(accel):
    big = scene(31, 120, 4, 96, 64, 6, sky=True)
    small = scene(32, 10, 1, 40, 24, 4)
    nb = big.camera.width * big.camera.height
    bufs = [DeviceArray(nb * 16), DeviceArray(nb * 16), DeviceArray(nb * 4)]
    try:
        # call 1: frames 1-3 of the large scene; then 4-5 continuing its accumulation
        exp = reference_entry(big, [3, 2], accel, kahan=False)
        acc, img = noalloc(big, 1, 3, accel, False, bufs)
        assert np.array_equal(acc.view(np.uint32), exp[0][0].reshape(-1, 4).view(np.uint32))
        assert np.array_equal(img, exp[0][1].reshape(-1))
        acc, img = noalloc(big, 4, 2, accel, False, bufs)
        assert np.array_equal(acc.view(np.uint32), exp[1][0].reshape(-1, 4).view(np.uint32))
        assert np.array_equal(img, exp[1][1].reshape(-1))
        # call 2: a smaller viewport and scene in the same buffers, Kahan
        exp = reference_entry(small, [2], accel, kahan=True)
        acc, img = noalloc(small, 1, 2, accel, True, bufs)
        assert np.array_equal(acc.view(np.uint32), exp[0][0].reshape(-1, 4).view(np.uint32))
        assert np.array_equal(img, exp[0][1].reshape(-1))
    finally:
        for b in bufs:
            b.free()
