"""Pins of the oracle's multi-bounce path and random stream
(Renderer.cu:300-384, BRDF.cu:72-116, Random.h:59-70).

The oracle logs, per bounce, the seed after `seed += i`, the light index,
Russian roulette's p and u and the direction it traced (ora_path_debug).
These are checked against

* tests/_pathref.py, a separate float64 restatement with Python-integer PCG:
  ids, seeds, light indices and u must be equal, p, directions and radiance
  equal to rounding, on paths of up to 6 bounces through diffuse and
  metallic spheres with two lights and the sky;
* closed forms that need no restatement at all: an albedo-1 convex sphere
  under the sky returns exactly (0.6, 0.7, 0.9) for every pixel; at albedo
  0.5 a pixel returns (0.6, 0.7, 0.9) * 0.5 / p or black, decided by the
  roulette draw computed from the hash chain; the first continuation of a
  metallic hit has cos(theta) = sqrt((1 - u1) / (1 + (a^2 - 1) u1)) with
  u1 from the hash chain, and a diffuse one sqrt(1 - u1).

A seed that advances by 1 instead of i, a dropped 1/p, a flipped roulette
comparison, the light picked from a different hash than the roulette draw,
or the samplers swapped each fail one of these.
"""
import math

import numpy as np
import pytest

import oracle as O
from ataraxia_amd._lib import LIGHT_DTYPE, MATERIAL_DTYPE, SPHERE_DTYPE

import _pathref as R

W, H = 24, 16


def pinhole_dirs(w, h, fov_deg=40.0):
    """Float32 primary directions of a camera at the origin looking down -z."""
    t = math.tan(math.radians(fov_deg) / 2)
    out = np.zeros((h, w, 3), np.float32)
    for y in range(h):
        for x in range(w):
            v = np.array([(2 * (x + 0.5) / w - 1) * t * w / h, (2 * (y + 0.5) / h - 1) * t, -1.0])
            out[y, x] = (v / np.linalg.norm(v)).astype(np.float32)
    return out


def build(spheres, mats, lights):
    sp = np.zeros(len(spheres), SPHERE_DTYPE)
    for i, (c, r, m) in enumerate(spheres):
        sp["center"][i], sp["radius"][i], sp["material"][i] = c, r, m
    ma = np.zeros(len(mats), MATERIAL_DTYPE)
    for i, m in enumerate(mats):
        ma["albedo"][i], ma["roughness"][i], ma["metallic"][i] = m["albedo"], m["roughness"], m["metallic"]
        ma["F0"][i], ma["emission_color"][i], ma["emission_intensity"][i] = m["F0"], m["emission"], m["intensity"]
    li = np.zeros(len(lights), LIGHT_DTYPE)
    for i, (p, c, k) in enumerate(lights):
        li["position"][i], li["color"][i], li["intensity"][i] = p, c, k
    # the float64 restatement reads the float32 values the oracle reads
    f = lambda v: tuple(float(np.float32(a)) for a in v)  # noqa: E731
    r_sph = [(f(c), float(np.float32(r)), m) for c, r, m in spheres]
    r_mat = [{"albedo": f(m["albedo"]), "roughness": float(np.float32(m["roughness"])),
              "metallic": float(np.float32(m["metallic"])), "F0": f(m["F0"]), "emission": f(m["emission"]),
              "intensity": float(np.float32(m["intensity"]))} for m in mats]
    r_li = [(f(p), f(c), float(np.float32(k))) for p, c, k in lights]
    return O.Scene(sp, ma, li), (r_sph, r_mat, r_li)


def mat(albedo, rough=0.5, metallic=0.0, emission=(0, 0, 0), intensity=0.0):
    return {"albedo": albedo, "roughness": rough, "metallic": metallic, "F0": (0.04, 0.04, 0.04),
            "emission": emission, "intensity": intensity}


# albedo lengths 0.87, 1.35, 0.61, 0.78, 1.03, 1.56: roulette with p < 1 and
# p = 1 both occur; the large sphere behind the cluster keeps paths going
CLUSTER = [((0.0, 0.0, -6.0), 1.2, 0), ((2.2, 0.3, -6.5), 1.0, 1), ((-2.1, -0.2, -5.8), 0.9, 2),
           ((0.0, 2.3, -6.2), 0.8, 3), ((0.3, -2.4, -6.0), 1.1, 4), ((0.0, 0.0, -14.0), 6.0, 5)]
CLUSTER_MATS = [mat((0.55, 0.5, 0.45), 0.7), mat((0.9, 0.8, 0.6), 0.3, 1.0), mat((0.3, 0.35, 0.4), 0.9),
                mat((0.45, 0.45, 0.45), 0.6, 1.0), mat((0.3, 0.9, 0.4), 0.2),
                mat((0.9, 0.9, 0.9), 0.8, emission=(1.0, 0.8, 0.6), intensity=0.5)]
CLUSTER_LIGHTS = [((3.0, 5.0, 0.0), (1.0, 0.95, 0.9), 10.0), ((-4.0, 2.0, -2.0), (0.6, 0.7, 1.0), 6.0)]


def test_multibounce_paths_match_float64_restatement():
    bounces = 6
    osc, (rs, rm, rl) = build(CLUSTER, CLUSTER_MATS, CLUSTER_LIGHTS)
    st = O.settings(bounces, sky_light=True)
    dirs = pinhole_dirs(W, H)
    origin = np.zeros(3, np.float32)
    checked = ambiguous = 0
    deep3 = deep = rr_end = rr_pass_below_one = ggx = cosine = 0
    frames = (1, 2, 7)
    for frame, y, x in ((f, y, x) for f in frames for y in range(H) for x in range(W)):
        if True:
            got = O.per_pixel_trace(x, y, W, dirs[y, x], origin, osc, frame, st)
            try:
                exp = R.per_pixel(x, y, W, dirs[y, x].astype(np.float64), origin, rs, rm, rl, frame, bounces, True)
            except R.Ambiguous:
                ambiguous += 1
                continue
            checked += 1
            n = len(exp["hit"])
            where = (x, y, frame)
            assert got["bounces"] == n, where
            assert got["hit"][:n].tolist() == exp["hit"] and np.all(got["hit"][n:] == -2), where
            assert got["seed"][:n].tolist() == exp["seed"], where
            ns = len(exp["shadow"])
            assert got["shadow"][:ns].tolist() == exp["shadow"], where
            assert got["light"][:len(exp["light"])].tolist() == exp["light"], where
            nr = len(exp["u"])
            assert np.array_equal(got["rr"][:nr, 1], np.float32(exp["u"])), where
            assert np.all(np.isnan(got["rr"][nr:])), where
            for b in range(n):
                tol = R.eps(b)
                np.testing.assert_allclose(got["dir"][b], exp["dir"][b], rtol=0, atol=tol, err_msg=str(where))
                if b < nr:
                    np.testing.assert_allclose(got["rr"][b, 0], exp["p"][b], rtol=tol, err_msg=str(where))
            # shading is as ill-conditioned as the geometry it reads (N.L of a
            # grazing light), hence ten times the direction tolerance
            np.testing.assert_allclose(got["rgba"][:3], exp["rgba"], rtol=10 * R.eps(n), atol=1e-6,
                                       err_msg=str(where))
            deep3 += n >= 3
            deep += n >= 4
            rr_end += any(u > p for u, p in zip(exp["u"], exp["p"]))
            rr_pass_below_one += any(u <= p < 1.0 for u, p in zip(exp["u"], exp["p"]))
            ggx += sum(s[0] == "ggx" for s in exp["sampler"])
            cosine += sum(s[0] == "cosine" for s in exp["sampler"])
    total = W * H * len(frames)
    assert ambiguous <= 0.2 * total, (ambiguous, total)
    # the comparison covered every branch of the loop
    assert deep3 >= 30 and deep >= 5 and rr_end >= 10 and rr_pass_below_one >= 10 and ggx >= 10 \
        and cosine >= 10, (deep3, deep, rr_end, rr_pass_below_one, ggx, cosine)


def _convex_scene(albedo, metallic=0.0, rough=0.5):
    return build([((0.0, 0.0, -5.0), 1.5, 0)], [mat(albedo, rough, metallic)], [])


def test_albedo_one_sphere_under_sky_is_exactly_sky():
    """p = 1, no roulette; a continuation leaves a convex sphere and sees the
    sky, so every pixel is (0.6, 0.7, 0.9) from two bounces on, and a hit
    pixel is black with one bounce."""
    osc, _ = _convex_scene((1.0, 1.0, 1.0))
    dirs = pinhole_dirs(W, H, 30.0)
    sky = np.float32([0.6, 0.7, 0.9])
    hits = 0
    for bounces in (1, 2, 5):
        for y in range(H):
            for x in range(W):
                g = O.per_pixel_trace(x, y, W, dirs[y, x], np.zeros(3, np.float32), osc, 3, O.settings(bounces, True))
                if g["hit"][0] == 0 and bounces == 1:
                    assert np.array_equal(g["rgba"][:3], np.float32([0, 0, 0]))
                else:
                    assert np.array_equal(g["rgba"][:3], sky), (x, y, bounces, g["rgba"])
                if g["hit"][0] == 0 and bounces > 1:
                    hits += 1
                    assert g["hit"][1] == -1 and g["rr"][0, 0] == 1.0
    assert hits > 50


def test_albedo_half_sphere_roulette_and_inverse_p():
    """throughput 0.5 gives p = sqrt(0.75) in float32; the roulette draw is the
    first hash of (x + y*W)*frame (no lights, so no light pick before it);
    a surviving path returns sky * (0.5 / p)."""
    osc, _ = _convex_scene((0.5, 0.5, 0.5))
    dirs = pinhole_dirs(W, H, 30.0)
    p = np.sqrt(np.float32(0.75))
    thr = np.float32(0.5) / p
    survive_exp = np.float32([0.6, 0.7, 0.9]) * thr
    frame = 5
    survived = ended = 0
    for y in range(H):
        for x in range(W):
            g = O.per_pixel_trace(x, y, W, dirs[y, x], np.zeros(3, np.float32), osc, frame, O.settings(2, True))
            if g["hit"][0] != 0:
                continue
            s0 = ((x + y * W) * frame) & R.M32
            u = np.float32(R.u32_to_unit(R.pcg_hash(s0)))
            assert g["seed"][0] == s0 and g["rr"][0, 1] == u and g["rr"][0, 0] == p
            if u > p:
                ended += 1
                assert np.array_equal(g["rgba"][:3], np.float32([0, 0, 0])) and g["bounces"] == 1
            else:
                survived += 1
                # bounce 1 starts from hash^3(s0) + 1: roulette, then u1 and u2
                assert g["seed"][1] == (R.pcg_hash(R.pcg_hash(R.pcg_hash(s0))) + 1) & R.M32
                assert np.array_equal(g["rgba"][:3], survive_exp), (x, y, g["rgba"], survive_exp)
    assert survived > 20 and ended > 5, (survived, ended)


@pytest.mark.parametrize("metallic,rough", [(1.0, 0.5), (1.0, 0.15), (0.0, 0.5)])
def test_first_continuation_direction_from_u1(metallic, rough):
    osc, (rs, _, _) = _convex_scene((1.0, 1.0, 1.0), metallic, rough)
    dirs = pinhole_dirs(W, H, 30.0)
    frame = 2
    checked = 0
    for y in range(0, H, 2):
        for x in range(0, W, 2):
            g = O.per_pixel_trace(x, y, W, dirs[y, x], np.zeros(3, np.float32), osc, frame, O.settings(2, True))
            if g["hit"][0] != 0:
                continue
            hit = R.trace((0.0, 0.0, 0.0), tuple(float(v) for v in dirs[y, x]), [(c, r) for c, r, _ in rs])
            N = np.array(hit[3])
            s0 = ((x + y * W) * frame) & R.M32
            u1 = R.u32_to_unit(R.pcg_hash(R.pcg_hash(s0)))
            if metallic > 0:
                a = rough * rough
                cos_exp = math.sqrt((1 - u1) / (1 + (a * a - 1) * u1))
            else:
                cos_exp = math.sqrt(1 - u1)
            d1 = g["dir"][1].astype(np.float64)
            assert abs(float(d1 @ N) - cos_exp) < 2e-5, (x, y, float(d1 @ N), cos_exp)
            assert abs(np.linalg.norm(d1) - 1.0) < 1e-5
            checked += 1
    assert checked > 15
