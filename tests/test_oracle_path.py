"""CPU pins of the oracle's path functions: traceRay, Cook-Torrance, the
samplers, perPixel known answers and the accumulation/display step."""
import ctypes
import math

import numpy as np

import oracle as O
from ataraxia_amd._lib import LIGHT_DTYPE, MATERIAL_DTYPE, SPHERE_DTYPE


def spheres(*items):
    a = np.zeros(len(items), SPHERE_DTYPE)
    for i, (c, r) in enumerate(items):
        a["center"][i] = c
        a["radius"][i] = r
        a["material"][i] = 0
    return a


def trace(o, d, sph):
    o = np.float32(o)
    d = np.float32(d)
    t = ctypes.c_float()
    i = O.lib.ora_trace_ray(o.ctypes.data, d.ctypes.data, sph.ctypes.data if len(sph) else None, len(sph),
                            ctypes.byref(t))
    return i, t.value


def test_trace_closed_forms():
    d = [0.0, 0.0, -1.0]
    # head-on: sphere centre 10 ahead, radius 2 -> t = 8
    i, t = trace([0, 0, 0], d, spheres(([0, 0, -10], 2.0)))
    assert i == 0 and abs(t - 8.0) < 1e-5
    # behind the origin: no hit
    i, _ = trace([0, 0, 0], d, spheres(([0, 0, 10], 2.0)))
    assert i == -1
    # origin inside a sphere: the far root is never used (t = min(t0,t1) = t0 < 0)
    i, _ = trace([0, 0, 0], d, spheres(([0, 0, -1], 5.0)))
    assert i == -1
    # offset just inside/outside the radius
    i, t = trace([1.99, 0, 0], d, spheres(([0, 0, -10], 2.0)))
    assert i == 0 and abs(t - (10 - math.sqrt(4 - 1.99 ** 2))) < 1e-4
    i, _ = trace([2.01, 0, 0], d, spheres(([0, 0, -10], 2.0)))
    assert i == -1
    # nearest of several, independent of order
    s = spheres(([0, 0, -30], 1.0), ([0, 0, -10], 1.0), ([0, 0, -20], 1.0))
    i, t = trace([0, 0, 0], d, s)
    assert i == 1 and abs(t - 9.0) < 1e-5
    # exact tie: identical spheres -> the lowest index wins (strict <)
    s = spheres(([0, 0, -30], 1.0), ([0, 0, -10], 1.0), ([0, 0, -10], 1.0))
    assert trace([0, 0, 0], d, s)[0] == 1
    # miss returns FLT_MAX
    i, t = trace([0, 0, 0], d, spheres())
    assert i == -1 and t == np.finfo(np.float32).max


def test_trace_against_float64_brute_force():
    rng = np.random.default_rng(11)
    for trial in range(40):
        n = 64
        s = np.zeros(n, SPHERE_DTYPE)
        s["center"] = rng.uniform(-10, 10, (n, 3))
        s["radius"] = rng.uniform(0.2, 2.0, n)
        for _ in range(20):
            o = rng.uniform(-12, 12, 3).astype(np.float32)
            d = rng.normal(size=3)
            d = (d / np.linalg.norm(d)).astype(np.float32)
            i, t = trace(o, d, s)
            # float64 closest positive near root
            oc = o.astype(np.float64) - s["center"].astype(np.float64)
            dd = d.astype(np.float64)
            b = oc @ dd
            c = (oc * oc).sum(1) - s["radius"].astype(np.float64) ** 2
            disc = b * b - (dd @ dd) * c
            tt = np.where(disc >= 0, (-b - np.sqrt(np.maximum(disc, 0))) / (dd @ dd), np.inf)
            tt = np.where(tt > 0, tt, np.inf)
            j = int(np.argmin(tt))
            if not np.isfinite(tt[j]):
                assert i == -1
                continue
            srt = np.sort(tt)
            if srt[1] - srt[0] > 1e-3 * srt[0]:      # not a near tie
                assert i == j
                assert abs(t - tt[j]) <= 1e-4 * tt[j] + 1e-5


def cook(albedo, F0, metallic, rough, N, V, L):
    out = np.zeros(3, np.float32)
    a = [np.float32(x) for x in (albedo, F0, N, V, L)]
    O.lib.ora_cook_torrance(a[0].ctypes.data, a[1].ctypes.data, metallic, rough, a[2].ctypes.data,
                            a[3].ctypes.data, a[4].ctypes.data, out.ctypes.data)
    return out


def test_cook_torrance_corners():
    N = np.float32([0, 1, 0])
    # V = L = N: VdotH = 1 -> Fresnel = F0 exactly; NdotH = 1 -> D = 1/(pi a2);
    # G = G1(1)^2 = 1; so f = ((1-F0)^2 (1-m) albedo/pi + D F0/(4+0.001)) * 1
    for rough in (0.25, 0.6, 1.0):
        for m in (0.0, 1.0):
            F0 = np.float32([0.04, 0.5, 0.9])
            alb = np.float32([0.8, 0.3, 0.1])
            f = cook(alb, F0, m, rough, N, N, N).astype(np.float64)
            a2 = float(rough) ** 4
            D = 1.0 / (math.pi * a2)
            F0d = F0.astype(np.float64)
            exp = (1 - F0d) * (1 - m) * (1 - F0d) * alb / math.pi + D * F0d / (4.0 + 0.001)
            np.testing.assert_allclose(f, exp, rtol=2e-6)
    # grazing view (VdotH -> 0 handled by the 1e-7 floors): finite and >= 0
    f = cook([1, 1, 1], [0.04] * 3, 0.0, 0.5, N, np.float32([1, 0, 0]), np.float32([0, 1, 0]))
    assert np.all(np.isfinite(f)) and np.all(f >= 0)
    # light below the horizon still contributes a little (NdotL floored at 1e-7)
    L = np.float32([0.894427, -0.447214, 0])
    f = cook([1, 1, 1], [0.04] * 3, 0.0, 0.5, N, N, L)
    assert np.all(f > 0) and np.all(f < 1e-5)
    # V exactly opposite L: H = normalize(0) is NaN, as in the reference
    f = cook([1, 1, 1], [0.04] * 3, 0.0, 0.5, N, N, np.float32([0, -1, 0]))
    assert np.all(np.isnan(f))


def test_ggx_distribution_normalisation():
    # integral over the hemisphere of D(h) cos(theta_h) dw = 1 for the GGX NDF
    for rough in (0.3, 0.6, 0.9):
        a2 = rough ** 4
        th = np.linspace(0, math.pi / 2, 200001)
        c = np.cos(th)
        D = a2 / (math.pi * ((c * c) * (a2 - 1) + 1) ** 2)
        integ = np.trapezoid(D * c * np.sin(th), th) * 2 * math.pi
        assert abs(integ - 1) < 1e-3
    # and the oracle's cookTorrance uses that D: specular lobe at NdotH=1 with
    # metallic=1 (no diffuse): f = D*G*F/(4 NdotL NdotV + 0.001) * NdotL
    N = np.float32([0, 0, 1])
    L = np.float32([0.6, 0, 0.8])
    V = np.float32([-0.6, 0, 0.8])          # H = N
    rough = 0.5
    f = cook([1, 1, 1], [1, 1, 1], 1.0, rough, N, V, L).astype(np.float64)
    a2 = rough ** 4
    D = 1 / (math.pi * a2)
    k = (rough + 1) ** 2 / 8
    G = (0.8 / (0.8 * (1 - k) + k)) ** 2
    exp = D * G * 1.0 / (4 * 0.8 * 0.8 + 0.001) * 0.8
    np.testing.assert_allclose(f, exp, rtol=1e-5)


def test_cosine_sampler_moments():
    rng = np.random.default_rng(4)
    N = np.float32([0.3, 0.8, -0.52])
    N = (N / np.linalg.norm(N)).astype(np.float32)
    zs = []
    out = np.zeros(3, np.float32)
    for s in rng.integers(0, 2 ** 32, 40000, dtype=np.uint64):
        st = ctypes.c_uint32(int(s))
        O.lib.ora_sample_cosine(N.ctypes.data, ctypes.byref(st), out.ctypes.data)
        zs.append(float(out @ N))
        assert abs(np.linalg.norm(out) - 1) < 1e-5
    z = np.array(zs)
    assert z.min() >= -1e-6
    assert abs(z.mean() - 2 / 3) < 0.006      # E[cos] = 2/3 under the cos/pi density
    assert abs((z * z).mean() - 0.5) < 0.006  # E[cos^2] = 1/2


def test_ggx_sampler_cdf():
    rng = np.random.default_rng(5)
    N = np.float32([0, 1, 0])
    out = np.zeros(3, np.float32)
    for rough in (0.3, 0.8):
        a2 = rough ** 4
        cs = []
        for s in rng.integers(0, 2 ** 32, 20000, dtype=np.uint64):
            st = ctypes.c_uint32(int(s))
            O.lib.ora_sample_ggx(N.ctypes.data, rough, ctypes.byref(st), out.ctypes.data)
            cs.append(float(out[1]))
        cs = np.sort(np.array(cs))
        # P(cos(theta_h) >= c) = (1 - c^2) / (1 + c^2 (a2 - 1)) for GGX D*cos sampling
        emp_ge = 1.0 - np.arange(1, len(cs) + 1) / len(cs)
        theo_ge = (1 - cs ** 2) / (1 + cs ** 2 * (a2 - 1))
        assert np.abs(emp_ge - theo_ge).max() < 0.015


def scene(sph, mats, lights):
    return O.Scene(sph, mats, lights)


def materials(n, albedo=(0.5, 0.5, 0.5), rough=0.5, metal=0.0, emis=(0, 0, 0), emis_i=0.0):
    m = np.zeros(n, MATERIAL_DTYPE)
    m["albedo"] = albedo
    m["roughness"] = rough
    m["metallic"] = metal
    m["F0"] = 0.04
    m["emission_color"] = emis
    m["emission_intensity"] = emis_i
    return m


def lights(*items):
    a = np.zeros(len(items), LIGHT_DTYPE)
    for i, (p, c, k) in enumerate(items):
        a["position"][i] = p
        a["color"][i] = c
        a["intensity"][i] = k
    return a


def pp(sc, bounces, sky=False, d=(0, 0, -1), o=(0, 0, 0), x=3, y=5, frame=1):
    rgba, hit, sh, rays = O.per_pixel(x, y, 16, np.float32(d), np.float32(o), sc, frame, O.settings(bounces, sky))
    return rgba, hit, sh, rays


def test_per_pixel_sky_and_zero_bounces():
    empty = scene(spheres(), materials(0), lights())
    rgba, hit, _, rays = pp(empty, 4, sky=True)
    assert np.array_equal(rgba, np.float32([0.6, 0.7, 0.9, 1.0])) and hit[0] == -1 and rays == 1
    rgba, _, _, _ = pp(empty, 4, sky=False)
    assert np.array_equal(rgba, np.float32([0, 0, 0, 1]))
    s = scene(spheres(([0, 0, -5], 1.0)), materials(1), lights(([0, 5, 0], (1, 1, 1), 10.0)))
    rgba, hit, _, rays = pp(s, 0, sky=True)
    assert np.array_equal(rgba, np.float32([0, 0, 0, 1])) and rays == 0


def test_per_pixel_emissive_one_bounce():
    s = scene(spheres(([0, 0, -5], 1.0)), materials(1, emis=(1.0, 0.5, 0.25), emis_i=4.0), lights())
    rgba, hit, _, _ = pp(s, 1)
    assert hit[0] == 0
    assert np.array_equal(rgba[:3], np.float32([4.0, 2.0, 1.0]))


def test_per_pixel_direct_light_closed_form():
    """One diffuse sphere, one light, one bounce: color = Le * f(N, V, L)
    evaluated in float64 from the geometry."""
    for light_pos in ([0.0, 5.0, -3.0], [4.0, 2.0, 1.0], [-3.0, -1.0, 2.0]):
        sph = spheres(([0, 0, -5], 1.0))
        m = materials(1, albedo=(0.7, 0.4, 0.2), rough=0.6)
        lig = lights((light_pos, (1.0, 0.9, 0.8), 12.0))
        rgba, hit, sh, rays = pp(scene(sph, m, lig), 1)
        assert hit[0] == 0 and sh[0] == -1 and rays == 2
        P = np.array([0, 0, -4.0])            # head-on hit
        Nn = np.array([0, 0, 1.0])
        V = np.array([0, 0, 1.0])
        L = np.array(light_pos) - P
        L /= np.linalg.norm(L)
        H = (V + L) / np.linalg.norm(V + L)
        ndl = max(Nn @ L, 1e-7)
        ndv = max(Nn @ V, 1e-7)
        ndh = max(Nn @ H, 0.0)
        vdh = max(V @ H, 0.0)
        F0 = 0.04
        F = F0 + (1 - F0) * (1 - min(max(vdh, 0), 1)) ** 5
        r = 0.6
        a2 = r ** 4
        D = a2 / (math.pi * (ndh ** 2 * (a2 - 1) + 1) ** 2)
        k = (r + 1) ** 2 / 8
        G = (ndv / (ndv * (1 - k) + k)) * (ndl / (ndl * (1 - k) + k))
        alb = np.array([0.7, 0.4, 0.2])
        f = ((1 - F) * (1 - F) * alb / math.pi + D * G * F / (4 * ndl * ndv + 0.001)) * ndl
        exp = np.array([1.0, 0.9, 0.8]) * 12.0 * f
        np.testing.assert_allclose(rgba[:3].astype(np.float64), exp, rtol=2e-5)


def test_per_pixel_occluded_light_and_no_lights():
    sph = spheres(([0, 0, -5], 1.0), ([0, 3, -4], 0.5))    # blocker between hit point and light
    m = materials(2)
    lig = lights(([0, 8, -4], (1, 1, 1), 10.0))
    rgba, hit, sh, _ = pp(scene(sph, m, lig), 1)
    assert hit[0] == 0 and sh[0] == 1
    assert np.array_equal(rgba[:3], np.float32([0, 0, 0]))
    # a blocker beyond the light does not occlude
    sph2 = spheres(([0, 0, -5], 1.0), ([0, 20, -4], 0.5))
    rgba, _, sh, _ = pp(scene(sph2, m, lig), 1)
    assert sh[0] == 1 and np.all(rgba[:3] > 0)
    # no lights: no shadow ray, no direct light
    rgba, hit, sh, rays = pp(scene(sph, m, lights()), 1)
    assert sh[0] == -2 and rays == 1 and np.array_equal(rgba[:3], np.float32([0, 0, 0]))


def test_seed_schedule_picks_light_from_first_hash():
    """Light index = PcgHash(seed) % L with seed = (x + y*W)*frame at bounce 0."""
    sph = spheres(([0, 0, -5], 1.0))
    m = materials(1)
    # light 0 is visible, light 1 sits inside the sphere (always occluded)
    lig = lights(([0, 5, 0], (1, 1, 1), 10.0), ([0, 0, -5], (1, 1, 1), 10.0))
    for x in range(6):
        seed = ((x + 5 * 16) * 3) & 0xFFFFFFFF
        li = O.pcg_hash(seed) % 2
        rgba, hit, sh, _ = pp(scene(sph, m, lig), 1, x=x, y=5, frame=3)
        lit = bool(np.any(rgba[:3] > 0))
        assert lit == (li == 0), (x, li, rgba)


def test_rgba8_truncation_table():
    cases = [((0, 0, 0, 0), 0), ((1, 1, 1, 1), 0xFFFFFFFF), ((0.5, 0, 0, 1), 0xFF00007F),
             ((0.999, 0.0039, 0.00393, 1), 0xFF0100FE), ((0, 0.2, 0, 1), 0xFF003300)]
    for rgba, packed in cases:
        v = np.float32(rgba)
        assert O.lib.ora_rgba8(v.ctypes.data) == packed, (rgba, hex(O.lib.ora_rgba8(v.ctypes.data)))
    v = np.float32([np.nan, 0, 0, 1])
    assert O.lib.ora_rgba8(v.ctypes.data) == 0xFF000000


def test_accumulation_mean_and_kahan():
    from ataraxia_amd import plain_cpu_scene
    from ataraxia_amd.scene import scene_from_camera_overrides
    sc = scene_from_camera_overrides(plain_cpu_scene(), 24, 16)
    cam = O.camera_from(sc.camera)
    osc = O.Scene(sc.spheres, sc.materials, sc.lights)
    st = O.settings(5)
    n = 24 * 16
    acc = np.zeros((n, 4), np.float32)
    kacc = np.zeros((n, 4), np.float32)
    comp = np.zeros((n, 4), np.float32)
    s64 = np.zeros((n, 4), np.float64)
    naive_ref = np.zeros((n, 4), np.float32)
    for f in range(1, 9):
        acc, _, img, rad, _ = O.render_frame(cam, osc, st, f, accum=acc)
        kacc, comp, kimg, _, _ = O.render_frame(cam, osc, st, f, accum=kacc, comp=comp)
        s64 += rad
        naive_ref = naive_ref + rad                 # float32 running sum in frame order
        assert np.array_equal(acc, naive_ref)
        assert np.all(acc[:, 3] == f)
        mean = np.clip(acc / np.float32(f), 0, 1)
        exp_img = ((mean[:, 3] * 255).astype(np.uint32) << 24) | ((mean[:, 2] * 255).astype(np.uint32) << 16) \
            | ((mean[:, 1] * 255).astype(np.uint32) << 8) | (mean[:, 0] * 255).astype(np.uint32)
        assert np.array_equal(img, exp_img)
    # Kahan's sum is at least as close to the float64 sum as the naive one
    assert np.abs(kacc - s64).sum() <= np.abs(acc - s64).sum()


def test_accumulate_pixels_equals_frame_by_frame():
    """The per-pixel accumulation the workload tests compare against equals
    the frame-by-frame oracle (naive and Kahan, accumulation on and off)."""
    from ataraxia_amd import plain_cpu_scene
    from ataraxia_amd.scene import scene_from_camera_overrides
    sc = scene_from_camera_overrides(plain_cpu_scene(), 20, 12)
    cam = O.camera_from(sc.camera)
    osc = O.Scene(sc.spheres, sc.materials, sc.lights)
    n = 20 * 12
    pix = np.arange(n)[::3]
    for kahan in (False, True):
        st = O.settings(5, False, True)
        acc = np.zeros((n, 4), np.float32)
        comp = np.zeros((n, 4), np.float32) if kahan else None
        for f in range(1, 6):
            acc, comp, img, _, _ = O.render_frame(cam, osc, st, f, accum=acc, comp=comp)
        got, gimg = O.accumulate_pixels(cam, osc, st, 1, 5, pix, kahan=kahan, threads=3)
        assert np.array_equal(got, acc[pix]) and np.array_equal(gimg, img[pix])
        # the float64 sum of the same float32 samples
        s64 = np.zeros((n, 4), np.float64)
        for f in range(1, 6):
            s64 += O.render_frame(cam, osc, st, f)[3]
        assert np.array_equal(O.accumulate_pixels_f64(cam, osc, st, 1, 5, pix, threads=2), s64[pix])
        # accumulation off: the index stays 1 and every frame restarts the sum
        st = O.settings(5, False, False)
        one, _, img1, _, _ = O.render_frame(cam, osc, st, 1, accum=np.zeros((n, 4), np.float32),
                                            comp=np.zeros((n, 4), np.float32) if kahan else None)
        got, gimg = O.accumulate_pixels(cam, osc, st, 1, 5, pix, kahan=kahan, threads=3)
        assert np.array_equal(got, one[pix]) and np.array_equal(gimg, img1[pix])
