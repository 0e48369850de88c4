"""CPU pins of the oracle's primitives: PCG, sin/cos/pow5, camera matrices."""
import json
import math
import os

import numpy as np

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M32 = 0xFFFFFFFF


def py_pcg(v):
    state = (v * 747796405 + 2891336453) & M32
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M32
    return ((word >> 22) ^ word) & M32


def np_pcg(v):
    v = v.astype(np.uint64)
    state = (v * np.uint64(747796405) + np.uint64(2891336453)) & np.uint64(M32)
    shift = (state >> np.uint64(28)) + np.uint64(4)
    word = (((state >> shift) ^ state) * np.uint64(277803737)) & np.uint64(M32)
    return ((word >> np.uint64(22)) ^ word) & np.uint64(M32)


def test_pcg_golden_vectors():
    g = json.load(open(os.path.join(GOLDEN, "pcg_vectors.json")))
    for v, h in zip(g["inputs"], g["outputs"]):
        assert O.pcg_hash(v) == h, v
    import ctypes
    for ch in g["chains"]:
        st = ctypes.c_uint32(ch["seed"])
        for expect in ch["states"]:
            u = O.lib.ora_pcg_float(ctypes.byref(st))
            assert st.value == expect
            # float(seed) / float(UINT32_MAX) == float32(seed) * 2^-32 exactly
            assert np.float32(u) == np.float32(np.float32(expect) * np.float32(2.0 ** -32))


def test_pcg_matches_python_and_is_a_permutation():
    rng = np.random.default_rng(0)
    xs = rng.integers(0, 2 ** 32, size=2000, dtype=np.uint64)
    for x in xs[:300]:
        assert O.pcg_hash(int(x)) == py_pcg(int(x))
    # PCG-RXS-M-XS is a bijection on u32: no collisions over a contiguous block
    block = np.arange(0, 1 << 20, dtype=np.uint64)
    h = np_pcg(block)
    assert len(np.unique(h)) == len(block)
    assert all(int(np_pcg(np.array([x]))[0]) == py_pcg(int(x)) for x in xs[:200])


def py_pcg_inverse(h):
    """Inverse of the hash (each PCG-RXS-M-XS step is invertible)."""
    word = h ^ (h >> 22)
    word = (word * pow(277803737, -1, 1 << 32)) & M32
    shift = (word >> 28) + 4          # the top 4 bits survive the xorshift
    state = word
    for _ in range(32 // shift + 1):  # undo state ^= state >> shift
        state = word ^ (state >> shift)
    return ((state - 2891336453) * pow(747796405, -1, 1 << 32)) & M32


def test_pcg_float_can_be_exactly_one():
    # float32(h) rounds up to 2^32 for h >= 2^32 - 128, so PcgFloat returns 1.0
    import ctypes
    for h in (0xFFFFFFFF, 0xFFFFFF80, 0xFFFFFF7F, 12345):
        seed = py_pcg_inverse(h)
        assert py_pcg(seed) == h and O.pcg_hash(seed) == h
        st = ctypes.c_uint32(seed)
        u = O.lib.ora_pcg_float(ctypes.byref(st))
        assert st.value == h
        assert (u == 1.0) == (h >= 0xFFFFFF80), (hex(h), u)


def ulp_err(got, exact):
    got = np.asarray(got, np.float32)
    e32 = np.asarray(exact, np.float64)
    spacing = np.spacing(np.abs(e32).astype(np.float32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - e32) / np.maximum(spacing, np.float64(2.0 ** -149))


def test_sin_cos_accuracy_on_path_domain():
    rng = np.random.default_rng(1)
    # the path evaluates sin/cos at (2*pi_f) * u with u = float(h) * 2^-32
    h = rng.integers(0, 2 ** 32, size=40000, dtype=np.uint64)
    u = (h.astype(np.float32) * np.float32(2.0 ** -32)).astype(np.float32)
    x = (np.float32(2.0) * np.float32(math.pi)).astype(np.float32) * u
    x = np.concatenate([x, np.float32([0.0, math.pi / 4, math.pi / 2, math.pi, 1.5 * math.pi, 2 * math.pi,
                                       6.2831855, 1e-8, 3.0e-5])]).astype(np.float32)
    s = np.array([O.lib.ora_sin(float(v)) for v in x], np.float32)
    c = np.array([O.lib.ora_cos(float(v)) for v in x], np.float32)
    xs = x.astype(np.float64)
    es, ec = np.sin(xs), np.cos(xs)
    # absolute error bound near zeros of sin/cos, ulp bound elsewhere
    big_s = np.abs(es) > 1e-3
    big_c = np.abs(ec) > 1e-3
    assert ulp_err(s[big_s], es[big_s]).max() <= 2.0
    assert ulp_err(c[big_c], ec[big_c]).max() <= 2.0
    assert np.abs(s - es).max() < 2e-7
    assert np.abs(c - ec).max() < 2e-7
    # a wrong quadrant or a dropped polynomial term would show up here
    assert abs(O.lib.ora_sin(1.5707964) - 1.0) < 1e-7 and abs(O.lib.ora_cos(3.1415927) + 1.0) < 1e-7


def test_pow5():
    rng = np.random.default_rng(2)
    x = rng.random(20000).astype(np.float32)
    got = np.array([O.lib.ora_pow5(float(v)) for v in x], np.float32)
    exact = x.astype(np.float64) ** 5
    ok = exact > 1e-30
    # three roundings (x^2, x^4, x^5): relative error <= 4*2^-24, i.e. <= 4 ulp
    err = ulp_err(got[ok], exact[ok])
    assert err.max() <= 4.0 and err.mean() < 1.0
    assert O.lib.ora_pow5(0.0) == 0.0 and O.lib.ora_pow5(1.0) == 1.0


def _mat(flat):
    return np.asarray(flat, np.float64).reshape(4, 4).T  # column-major -> row-major


def test_perspective_closed_form():
    for fov, aspect, n, f in [(45.0, 16 / 9, 0.1, 100.0), (90.0, 1.0, 0.5, 50.0), (30.0, 0.5, 1.0, 10.0)]:
        out = np.zeros(16, np.float32)
        O.lib.ora_perspective(fov, aspect, n, f, out.ctypes.data)
        t = math.tan(math.radians(fov) / 2)
        ref = np.zeros((4, 4))
        ref[0, 0] = 1 / (aspect * t)
        ref[1, 1] = 1 / t
        ref[2, 2] = -(f + n) / (f - n)
        ref[3, 2] = -1.0
        ref[2, 3] = -(2 * f * n) / (f - n)
        np.testing.assert_allclose(_mat(out), ref, rtol=2e-6, atol=1e-7)


def test_look_at_and_inverse():
    rng = np.random.default_rng(3)
    for _ in range(20):
        eye = rng.uniform(-10, 10, 3).astype(np.float32)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        d = d.astype(np.float32)
        center = (eye + d).astype(np.float32)
        up = np.float32([0, 1, 0])
        out = np.zeros(16, np.float32)
        O.lib.ora_look_at(eye.ctypes.data, center.ctypes.data, up.ctypes.data, out.ctypes.data)
        V = _mat(out)
        R = V[:3, :3]
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=2e-6)
        # the camera looks down -z: V maps eye to the origin and eye+d to (0,0,-1)
        np.testing.assert_allclose(V @ np.append(eye, 1), [0, 0, 0, 1], atol=2e-5)
        np.testing.assert_allclose(V @ np.append(center, 1), [0, 0, -1, 1], atol=2e-5)
        inv = np.zeros(16, np.float32)
        O.lib.ora_inverse(out.ctypes.data, inv.ctypes.data)
        np.testing.assert_allclose(_mat(inv) @ V, np.eye(4), atol=2e-5)
        np.testing.assert_allclose(_mat(inv), np.linalg.inv(V), atol=2e-5)
    # general matrices too (cofactor signs and indices)
    for _ in range(20):
        A = rng.normal(size=(4, 4)) + 4 * np.eye(4)
        flat = A.T.astype(np.float32).ravel()
        inv = np.zeros(16, np.float32)
        O.lib.ora_inverse(flat.ctypes.data, inv.ctypes.data)
        np.testing.assert_allclose(_mat(inv), np.linalg.inv(_mat(flat)), rtol=1e-4, atol=1e-5)


def test_ray_directions_invariants():
    cam = O.camera((0.0, 3.0, 12.0), (0.0, -0.24253562, -0.9701425), 64, 36, fov=45.0)
    d = O.ray_directions(cam).astype(np.float64)
    np.testing.assert_allclose(np.linalg.norm(d, axis=2), 1.0, atol=3e-7)
    fwd = np.array(cam.direction, np.float64)
    fwd /= np.linalg.norm(fwd)
    # no half-pixel offset: pixel (W/2, H/2) maps to clip (0,0), the view axis
    np.testing.assert_allclose(d[18, 32], fwd, atol=2e-6)
    # pixel (W/2, 0) lies on the bottom edge of the frustum: angle = fov/2
    ang = math.degrees(math.acos(np.clip(np.dot(d[0, 32], fwd), -1, 1)))
    assert abs(ang - 22.5) < 1e-3
    # pixel (0, H/2) lies on the left edge: tan(angle) = aspect * tan(fov/2)
    ang = math.atan(64 / 36 * math.tan(math.radians(22.5)))
    assert abs(math.acos(np.dot(d[18, 0], fwd)) - ang) < 1e-5
    assert d[18, 0][0] < 0 and d[0, 32][1] < d[18, 32][1]
