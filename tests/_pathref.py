"""Independent float64 restatement of Renderer::perPixel for small scenes.

Test infrastructure: a second, separately written statement of the path
that the C oracle is checked against, so that a slip shared by the oracle and
the kernels (a wrong seed increment, a dropped 1/p, a flipped roulette test,
the wrong continuation sampler) fails a test.  It is written from the
reference source, not from the oracle:

  Core/include/Random.h:59-70     PcgHash / PcgFloat, with Python integers
  Engine/src/Renderer.cu:251-285  traceRay (literal 2b / 4ac form)
  Engine/src/Renderer.cu:287-387  perPixel, every step logged per bounce
  Engine/src/Renderer.cu:396-409  rayHit
  Engine/src/BRDF.cu:9-70         cookTorrance in float64 (libm pow)
  Engine/src/BRDF.cu:72-116       cosine and GGX samplers (libm sin/cos)

Everything continuous is float64, so it differs from the float32 oracle by
rounding, and that difference grows along a path: float32 hit distances near
a sphere's silhouette are ill-conditioned (the discriminant cancels), and a
hit point that is off by d moves the next bounce's normal, origin and
direction by about d / radius.  Measured on the test scenes the directions
differ by ~1e-5 after one bounce and grow ~2-3x per bounce.  So every
discrete decision whose outcome rounding could change is measured against a
margin that grows with the bounce, eps(b) = 2e-5 * 3^b: a closest hit whose
t is within eps of another candidate's, a grazing ray (discriminant within
eps of zero relative to its terms), a hit at t within eps of zero, a shadow
test t^2 within eps of the light distance^2, a roulette u within eps of p,
and a tangent frame whose |N.x| and |N.y| are within eps.  A path with any
such decision is reported as ambiguous and is not compared; everything else
must agree exactly (ids, seeds, light indices, u) or to eps(b) (p,
directions, radiance).
"""
from __future__ import annotations

import math

import numpy as np

M32 = 0xFFFFFFFF


def eps(bounce: int) -> float:
    """Decision margin and comparison tolerance at a bounce."""
    return 2e-5 * 3.0 ** bounce


def pcg_hash(seed: int) -> int:
    state = (seed * 747796405 + 2891336453) & M32
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M32
    return ((word >> 22) ^ word) & M32


def u32_to_unit(v: int) -> float:
    # static_cast<float>(seed) rounds to nearest even; float(UINT32_MAX) is 2^32
    return float(np.float32(v)) / 4294967296.0


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def _scale(a, s):
    return (a[0] * s, a[1] * s, a[2] * s)


def _mul(a, b):
    return (a[0] * b[0], a[1] * b[1], a[2] * b[2])


def _norm(a):
    return _scale(a, 1.0 / math.sqrt(_dot(a, a)))


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


class Ambiguous(Exception):
    pass


def trace(o, d, spheres, e=2e-5):
    """Closest hit by the reference's loop; raises Ambiguous within the
    relative margin e of a decision boundary.  spheres: list of (center, radius)."""
    best, tmin = -1, math.inf
    cands = []
    for i, (c, r) in enumerate(spheres):
        oc = _sub(o, c)
        a = _dot(d, d)
        b = 2.0 * _dot(oc, d)
        cc = _dot(oc, oc) - r * r
        disc = b * b - 4.0 * a * cc
        scale = b * b + abs(4.0 * a * cc)
        if abs(disc) <= e * scale:
            # grazing: whether the rounded discriminant is negative is not known
            t_graze = -b / (2.0 * a)
            if t_graze > 0.0:
                raise Ambiguous("grazing ray")
            continue
        if disc < 0.0:
            continue
        t = (-b - math.sqrt(disc)) / (2.0 * a)
        if abs(t) <= e * (math.sqrt(_dot(oc, oc)) + r):
            raise Ambiguous("hit at t ~ 0")
        if t > 0.0:
            cands.append((t, i))
            if t < tmin:
                best, tmin = i, t
    for t, i in cands:
        if i != best and abs(t - tmin) <= e * tmin:
            raise Ambiguous("near tie")
    if best < 0:
        return None
    c = spheres[best][0]
    p = _add(_sub(o, c), _scale(d, tmin))
    n = _norm(p)
    return best, tmin, _add(p, c), n


def cook_torrance(albedo, F0, metallic, rough, N, V, L):
    H = _norm(_add(V, L))
    ndl = max(_dot(N, L), 1e-7)
    ndv = max(_dot(N, V), 1e-7)
    ndh = max(_dot(N, H), 0.0)
    vdh = max(_dot(V, H), 0.0)
    p5 = (1.0 - min(max(vdh, 0.0), 1.0)) ** 5
    F = tuple(f + (1.0 - f) * p5 for f in F0)
    a = rough * rough
    a2 = a * a
    den = ndh * ndh * (a2 - 1.0) + 1.0
    D = a2 / (math.pi * den * den)
    k = (rough + 1.0) ** 2 / 8.0
    G = (ndv / (ndv * (1.0 - k) + k)) * (ndl / (ndl * (1.0 - k) + k))
    spec = tuple(D * G * f / (4.0 * ndl * ndv + 0.001) for f in F)
    kd = tuple((1.0 - f) * (1.0 - metallic) for f in F)
    diff = tuple((1.0 - f) * al / math.pi for f, al in zip(F, albedo))
    return tuple((kd[j] * diff[j] + spec[j]) * ndl for j in range(3))


def _frame(N):
    if abs(N[0]) > abs(N[1]):
        T = _scale((-N[2], 0.0, N[0]), 1.0 / math.sqrt(N[0] * N[0] + N[2] * N[2]))
    else:
        T = _scale((0.0, -N[2], N[1]), 1.0 / math.sqrt(N[1] * N[1] + N[2] * N[2]))
    return T, _cross(N, T)


def _to_world(N, h, e):
    T, B = _frame(N)
    if abs(abs(N[0]) - abs(N[1])) <= e * (abs(N[0]) + abs(N[1])):
        raise Ambiguous("tangent frame choice")
    return _add(_add(_scale(T, h[0]), _scale(B, h[1])), _scale(N, h[2]))


def per_pixel(x, y, width, direction, origin, spheres, materials, lights, frame, max_bounces, sky):
    """One path.  materials: list of dicts (albedo, roughness, metallic, F0,
    emission, intensity) indexed by the sphere's material id; spheres: list of
    (center, radius, material).  Returns a dict with the radiance and the
    per-bounce log, or raises Ambiguous."""
    geo = [(c, r) for c, r, _ in spheres]
    ro = tuple(float(v) for v in origin)
    rd = tuple(float(v) for v in direction)
    color = (0.0, 0.0, 0.0)
    thr = (1.0, 1.0, 1.0)
    seed = ((x + y * width) * frame) & M32
    log = {"hit": [], "shadow": [], "seed": [], "light": [], "p": [], "u": [], "dir": [], "sampler": []}
    for i in range(max_bounces):
        e = eps(i)
        seed = (seed + i) & M32
        log["seed"].append(seed)
        log["dir"].append(rd)
        h = trace(ro, rd, geo, e)
        log["hit"].append(-1 if h is None else h[0])
        if h is None:
            if sky:
                color = _add(color, _mul((0.6, 0.7, 0.9), thr))
            break
        idx, t, P, N = h
        m = materials[spheres[idx][2]]
        if m["intensity"] > 0.0:
            color = _add(color, _mul(_scale(m["emission"], m["intensity"]), thr))
        F0m = tuple(f * (1.0 - m["metallic"]) + a * m["metallic"] for f, a in zip(m["F0"], m["albedo"]))
        if lights:
            li = pcg_hash(seed) % len(lights)
            log["light"].append(li)
            lp, lc, lk = lights[li]
            Lv = _sub(lp, P)
            dist2 = _dot(Lv, Lv)
            L = _norm(Lv)
            so = _add(P, _scale(N, 1e-4))
            sh = trace(so, L, geo, e)
            log["shadow"].append(-1 if sh is None else sh[0])
            if sh is not None and abs(sh[1] * sh[1] - dist2) <= e * dist2:
                raise Ambiguous("occluder at the light")
            if not (sh is not None and sh[1] * sh[1] < dist2):
                f = cook_torrance(m["albedo"], F0m, m["metallic"], m["roughness"], N, _scale(rd, -1.0), L)
                color = _add(color, _mul(_mul(_scale(lc, lk), f), thr))
        else:
            log["light"].append(-1)
            log["shadow"].append(-2)
        thr = _mul(thr, m["albedo"])
        ro = _add(P, _scale(N, 1e-4))
        p = max(0.1, min(1.0, math.sqrt(_dot(thr, thr))))
        seed = pcg_hash(seed)
        u = u32_to_unit(seed)
        log["p"].append(p)
        log["u"].append(u)
        if abs(u - p) <= e and p < 1.0:
            raise Ambiguous("roulette")
        if u > p:
            break
        thr = _scale(thr, 1.0 / p)
        seed = pcg_hash(seed)
        u1 = u32_to_unit(seed)
        seed = pcg_hash(seed)
        u2 = u32_to_unit(seed)
        if m["metallic"] > 0.0:
            a = m["roughness"] ** 2
            cos_t = math.sqrt((1.0 - u1) / (1.0 + (a * a - 1.0) * u1))
            sin_t = math.sqrt(max(0.0, 1.0 - cos_t * cos_t))
            phi = 2.0 * math.pi * u2
            hl = (sin_t * math.cos(phi), sin_t * math.sin(phi), cos_t)
            log["sampler"].append(("ggx", u1, u2, cos_t))
        else:
            r = math.sqrt(u1)
            th = 2.0 * math.pi * u2
            hl = (r * math.cos(th), r * math.sin(th), math.sqrt(1.0 - u1))
            log["sampler"].append(("cosine", u1, u2, hl[2]))
        rd = _to_world(N, hl, e)
    log["rgba"] = color
    return log
