// trace_device.h -- closest-hit sphere loops (Engine/src/Renderer.cu:251-285).
//
// The reference tests every sphere with
//   oc = o - c; a = d.d; b = 2 oc.d; c = oc.oc - r*r; disc = b*b - 4*a*c
//   t0 = (-b - sqrt(disc)) / (2a); t1 = (-b + sqrt(disc)) / (2a); t = min(t0,t1)
//   accept iff t > 0 && t < tmin   (strict <: the lowest index wins ties)
// Both loops below evaluate the reduced form
//   b' = oc.d,  disc' = b'*b' - a*c,  t = (-b' - sqrt(disc')) / a
// which rounds to exactly the same t: b = 2b', disc = 4disc' and
// sqrt(4x) = 2sqrt(x) are exact power-of-two scalings in binary floating
// point (barring overflow and results in the subnormal range), so every
// intermediate is the reference's divided by a power of two.  t1 >= t0 always,
// so min(t0, t1) is t0 and t1 need not be formed.  17 FLOP per test.
//
// trace_exact follows the reference literally: correctly rounded sqrt and
// division for every sphere with disc >= 0.
//
// trace_filtered gets the same answer with one hardware square root
// (v_sqrt_f32) per sphere and no division.  It keeps, per lane, an interval
// [lo, hi] around the approximate numerator n~ = -b' - sqrt~(disc') that is
// guaranteed to contain the exact numerator n = RN(-b' - RN(sqrt(disc'))):
//   |n~ - n| <= 2^-20 (s~ + |n~|)   (v_sqrt_f32 within 4 ulp, two roundings)
// and widens it to 2^-17 (s~ + |n~|) so that two intervals that do not overlap
// order the exact quotients n/a strictly (relative gap >= 2^-18 >> 1 ulp).
// Because n > 0 <=> t > 0 (a = d.d is within a few ulp of 1) and division by
// a > 0 is monotone, a candidate whose interval lies wholly below the current
// best and wholly above zero is the reference's new best, and a candidate
// whose interval lies wholly at/above the best or at/below zero is rejected by
// the reference too.  Anything else (overlapping intervals: two spheres within
// ~1e-5 relative distance along the ray; a sign that the interval cannot
// decide) marks the lane ambiguous and that lane re-runs trace_exact.  Where
// v_sqrt_f32 is not trusted (disc' below 2^-96) the interval's half-width is
// floored at 2^-47, above any error of the square root there.  Without ambiguity the winner's
// exact t is recomputed once with the exact formula.
#pragma once

#include <float.h>
#include "path_math.h"

namespace atx {

struct Hit {
    float t;     // tmin (FLT_MAX on a miss)
    int32_t id;  // closest sphere or -1
};

template <class Src>
__device__ __forceinline__ Hit trace_exact(const Src& src, uint32_t n, f3 o, f3 d)
{
    const float a = dot(d, d);
    float tmin = FLT_MAX;
    int32_t best = -1;
    for (uint32_t i = 0; i < n; ++i) {
        const float4 s = src(i);
        const float ocx = o.x - s.x, ocy = o.y - s.y, ocz = o.z - s.z;
        const float b = (ocx * d.x + ocy * d.y) + ocz * d.z;
        const float c = ((ocx * ocx + ocy * ocy) + ocz * ocz) - s.w;
        const float disc = b * b - a * c;
        if (disc >= 0.0f) {
            const float t = (-b - sqrt_cr(disc)) / a;
            if (t > 0.0f && t < tmin) {
                tmin = t;
                best = (int32_t)i;
            }
        }
    }
    Hit h;
    h.t = tmin;
    h.id = best;
    return h;
}

// Exact t of one sphere (same expression as trace_exact's loop body).
__device__ __forceinline__ float exact_t(float4 s, f3 o, f3 d, float a)
{
    const float ocx = o.x - s.x, ocy = o.y - s.y, ocz = o.z - s.z;
    const float b = (ocx * d.x + ocy * d.y) + ocz * d.z;
    const float c = ((ocx * ocx + ocy * ocy) + ocz * ocz) - s.w;
    const float disc = b * b - a * c;
    return (-b - sqrt_cr(disc)) / a;
}

// Running state of the filtered closest-hit search of one ray.
struct Filter {
    float best_lo;   // interval of the best numerator so far
    float best_hi;
    int32_t best;    // its sphere index, -1 = none
    uint32_t amb;    // nonzero: the filter could not decide, re-run the exact loop (an
                     // integer, so that loops carry it in a VGPR, not a lane mask)
};

__device__ __forceinline__ Filter filter_init()
{
    Filter f;
    f.best_lo = __builtin_inff();
    f.best_hi = __builtin_inff();
    f.best = -1;
    f.amb = 0u;
    return f;
}

// One filtered test of sphere s (index id) against the ray (o, d), a = d.d.
// The decision does not depend on the order in which spheres are visited:
// "better" means a strictly smaller exact t, "near" an interval that
// overlaps the best one.  Returns true when s became the new best.
__device__ __forceinline__ bool filter_sphere(float4 s, int32_t id, f3 o, f3 d, float a, Filter& f)
{
    constexpr float kMargin = 0x1p-17f;
    const float ocx = o.x - s.x, ocy = o.y - s.y, ocz = o.z - s.z;
    const float b = (ocx * d.x + ocy * d.y) + ocz * d.z;
    const float c = ((ocx * ocx + ocy * ocy) + ocz * ocz) - s.w;
    const float disc = b * b - a * c;
    // No lane's line meets the sphere (disc < 0 or NaN): the bounds below
    // would be NaN and change nothing, so the wave skips them.
    if (!__any(disc >= 0.0f))
        return false;
    const float sq = __builtin_amdgcn_sqrtf(disc);  // NaN when disc < 0
    const float nn = -b - sq;
    // The floor keeps the interval sound where v_sqrt_f32 is not trusted
    // (disc < 2^-96, subnormal inputs): the half-width is then >= 2^-47,
    // above any error of a square root below 2^-48 plus the subtraction's
    // rounding.  (disc < 0 gives NaN bounds, which every comparison rejects.)
    const float u = __builtin_fmaxf(sq + __builtin_fabsf(nn), 0x1p-30f);
    const float hi = __builtin_fmaf(u, kMargin, nn);
    const float lo = __builtin_fmaf(u, -kMargin, nn);
    const bool better = (hi < f.best_lo) & (lo > 0.0f);
    const bool near = (hi > 0.0f) & (lo < f.best_hi);
    f.amb |= (near & !better) ? 1u : 0u;
    f.best_lo = better ? lo : f.best_lo;
    f.best_hi = better ? hi : f.best_hi;
    f.best = better ? id : f.best;
    return better;
}

template <class Src>
__device__ __forceinline__ Hit trace_filtered(const Src& src, uint32_t n, f3 o, f3 d, uint32_t& exact_count)
{
    const float a = dot(d, d);
    Filter f = filter_init();
    for (uint32_t i = 0; i < n; ++i)
        filter_sphere(src(i), (int32_t)i, o, d, a, f);
    const bool amb = f.amb != 0u;
    const int32_t best = f.best;
    Hit h;
    if (amb) {
        ++exact_count;
        h = trace_exact(src, n, o, d);
    } else {
        h.id = best;
        h.t = best >= 0 ? exact_t(src((uint32_t)best), o, d, a) : FLT_MAX;
    }
    return h;
}

}  // namespace atx
