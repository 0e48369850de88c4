// trace_bvh.h -- closest-hit and shadow queries through the sphere BVH with
// the flat loop's exact result (Engine/src/Renderer.cu:251-285 semantics).
//
// Which spheres a box may hide.  Let D bound |o - c| for the tree's spheres
// (the ray origin o to a sphere centre c) and f(t) = |o + t d - c|^2 - r^2.
// The reference's rounded evaluation (oc, b, c, disc, sqrt, numerator and
// quotient each rounded once, u = 2^-24) returns a t~ with
//   |f(t~)| <= 93 u D^2 + 36 u r^2 < 2^-17 (D^2 + r^2)
// (forward error of b within 4.01 u D, of c within 6.02 u D^2 + u r^2; the
// disc rounding moves c by at most 2.02 u (2 D^2 + r^2); the root of the
// perturbed quadratic then satisfies |f| <= 48.2 u D^2 + 16.1 u r^2, and the
// last three roundings move t by at most u (5.03 D + 3.02 r) on a slope of at
// most 6 D + 2 r).  So every hit the flat loop can report lies in the sphere
// of radius sqrt(r^2 + 2^-17 (D^2 + r^2)): the box inflation per ray is the
// largest excess of that radius over r across the tree's radius range
// (attained at r_min or r_max), plus a rounding term.  Node boxes are 16-bit
// cells of the tree's box (world = base + cell * scale); the ray is moved
// into cell units (o_cell = (o - base) / scale, 1/d scaled by scale), and
// the rounding term, 2^-18 (|o - base|_inf + 65535 scale_max + m), covers
// that transform and the slab test (at most ~6 u of those magnitudes).
//
// Culling by distance uses an upper bound of the best sphere's exact t
// (best_hi / a, rounded up by 2^-20): a box whose inflated entry lies beyond
// it holds only spheres with a strictly larger exact t, which the reference
// rejects whatever their index.  Shadow queries (dist2 >= 0: only the
// reference's decision t > 0 && t*t < dist2 on the closest hit matters) also
// skip boxes beyond sqrt(dist2), where no sphere can satisfy t*t < dist2, and
// stop at the first sphere whose exact t is proven below sqrt(dist2); the
// sphere they return is then some occluder, not necessarily the closest.
//
// Spheres that are visited go through filter_sphere, like the flat loop; a
// lane whose filter is ambiguous re-runs the exact flat loop, and origins so
// far away that the inflation cannot be formed use the flat filtered loop.
#pragma once

#include "bvh.h"
#include "trace_device.h"

namespace atx {

#if defined(__HIP_DEVICE_COMPILE__)
#define ATX_GLOBAL __attribute__((address_space(1)))
#else
#define ATX_GLOBAL
#endif

// Spheres found in the tree's leaves are named by their leaf slot (tagged)
// while the query runs and mapped to their scene index once at its end: the
// filter's decisions do not depend on sphere indices (ties go to the exact
// loop), and the index load then stays off the leaf loop's critical path.
// A scene with a tree has at most kBvhMaxTree + kBvhMaxBig (16398) spheres,
// so scene indices never carry the tag.
constexpr int32_t kLeafSlot = 0x40000000;

struct BvhQuery {
    uint32_t tests = 0;   // sphere tests
    uint32_t boxes = 0;   // node box tests
    uint32_t exact = 0;   // exact fallback loops
};

__device__ __forceinline__ float bvh_inv_dir(float v)
{
    const float tiny = 0x1p-64f;
    const float c = __builtin_fabsf(v) < tiny ? __builtin_copysignf(tiny, v) : v;
    return __builtin_amdgcn_rcpf(c);
}

// min(a, b) as one v_min_f32.  fminf on a value carried around the node loop
// (tlim) costs a canonicalising v_max_f32 per node besides the min; neither
// operand can be a signalling NaN here (both come out of arithmetic), for
// which that canonicalisation exists.
__device__ __forceinline__ float bvh_min_raw(float a, float b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    float r;
    asm("v_min_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
#else
    return __builtin_fminf(a, b);
#endif
}

// Radius excess sqrt(r^2 + q) - r, q = 2^-17 (D^2 + r^2), in a form without
// cancellation; hardware sqrt and reciprocal (a few ulp, which the caller's
// 2^-16 widening covers).
__device__ __forceinline__ float bvh_excess(float r, float D2)
{
    const float q = 0x1p-17f * (D2 + r * r);
    return q * __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(r * r + q) + r);
}

// Closest hit (dist2 < 0) or shadow query (dist2 >= 0) of ray (o, d).
// `flat` reads sphere i of the scene in its original order (n_all spheres).
// kPriority: raise the wave's priority in the node loop (below).
template <bool kPriority = false, class FlatSrc>
__device__ __forceinline__ Hit trace_bvh(const BvhDev& B, const FlatSrc& flat, uint32_t n_all, f3 o, f3 d,
                                         float dist2, BvhQuery& q)
{
    const float a = dot(d, d);
    const bool shadow = dist2 >= 0.0f;
    Filter f = filter_init();

    // spheres outside the tree (wave-uniform indices)
    for (uint32_t j = 0; j < B.n_big; ++j) {
        const int32_t id = B.big[j];
        filter_sphere(flat((uint32_t)id), id, o, d, a, f);
    }
    q.tests += B.n_big;

    // per-ray box inflation (world units), then the ray in cell units
    const float ox = o.x - B.cx, oy = o.y - B.cy, oz = o.z - B.cz;
    const float dist = __builtin_amdgcn_sqrtf((ox * ox + oy * oy) + oz * oz);
    const float D = (dist * (1.0f + 0x1p-18f) + B.rb) * (1.0f + 0x1p-20f);
    const float D2 = D * D;
    const float m = __builtin_fmaxf(bvh_excess(B.r_min, D2), bvh_excess(B.r_max, D2)) * (1.0f + 0x1p-16f);
    const float rx = o.x - B.bx, ry = o.y - B.by, rz = o.z - B.bz;
    const float omax = __builtin_fmaxf(__builtin_fmaxf(__builtin_fabsf(rx), __builtin_fabsf(ry)), __builtin_fabsf(rz));
    const float e = 0x1p-18f * ((omax + B.ext) + m) + 0x1p-60f;
    const float M = (m + e) * (1.0f + 0x1p-20f);
    if (!(M < 0x1p60f) || !(omax < 0x1p60f)) {
        // origin too far for the inflation to be formed: flat loop over everything
        q.tests += n_all;
        uint32_t ex = 0;
        const Hit h = trace_filtered(flat, n_all, o, d, ex);
        q.exact += ex;
        return h;
    }
    // t = (cell - o_cell) * (scale / d) along each axis
    const float ix = bvh_inv_dir(d.x) * B.sx, iy = bvh_inv_dir(d.y) * B.sy, iz = bvh_inv_dir(d.z) * B.sz;
    const float gx = rx * B.isx, gy = ry * B.isy, gz = rz * B.isz;
    const float mx = M * B.isx * (1.0f + 0x1p-20f), my = M * B.isy * (1.0f + 0x1p-20f),
                mz = M * B.isz * (1.0f + 0x1p-20f);
    const float px = -((gx + mx) * ix), py = -((gy + my) * iy), pz = -((gz + mz) * iz);
    const float qx = -((gx - mx) * ix), qy = -((gy - my) * iy), qz = -((gz - mz) * iz);
    const float inv_a_up = __builtin_amdgcn_rcpf(a) * (1.0f + 0x1p-20f);
    const float s_up = shadow ? __builtin_amdgcn_sqrtf(dist2) * (1.0f + 0x1p-18f) : __builtin_inff();
    float tlim = __builtin_fminf(f.best_hi * inv_a_up, s_up);

    const uint32_t n_nodes = B.n_nodes;
    const uint32_t oct = ((d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u)) & B.layout_mask;
    const ATX_GLOBAL uint4* nodes = (const ATX_GLOBAL uint4*)(uintptr_t)B.nodes + oct * n_nodes;
    const ATX_GLOBAL float4* lsph = (const ATX_GLOBAL float4*)(uintptr_t)B.lsph;
    const ATX_GLOBAL int32_t* lidx = (const ATX_GLOBAL int32_t*)(uintptr_t)B.lidx;
    uint32_t i = (shadow && tlim * tlim < dist2) ? n_nodes : 0u;
    // Wave priority (kPriority, for the persistent path kernel, whose waves
    // do other work between traces): a wave in this loop issues before waves
    // outside it (job hand-out, shading), and the node load of each trip
    // before the rest of the loop's work, so that the loop's chain of
    // dependent node fetches waits less for issue slots (1080p path kernel:
    // 28.97 -> 27.12 ms per 256-frame launch, 3.15 -> 3.04 ms per 20-frame
    // one).
    if constexpr (kPriority)
        __builtin_amdgcn_s_setprio(2);
    while (i < n_nodes) {
        if constexpr (kPriority)
            __builtin_amdgcn_s_setprio(3);
        const uint4 w = nodes[i];
        if constexpr (kPriority)
            __builtin_amdgcn_s_setprio(2);
        ++q.boxes;
        const float lx = (float)(w.x & 0xffffu), ly = (float)(w.x >> 16), lz = (float)(w.y & 0xffffu);
        const float hx = (float)(w.y >> 16), hy = (float)(w.z & 0xffffu), hz = (float)(w.z >> 16);
        const float ax = __builtin_fmaf(lx, ix, px), bx = __builtin_fmaf(hx, ix, qx);
        const float ay = __builtin_fmaf(ly, iy, py), by = __builtin_fmaf(hy, iy, qy);
        const float az = __builtin_fmaf(lz, iz, pz), bz = __builtin_fmaf(hz, iz, qz);
        const float tn = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(ax, bx), __builtin_fminf(ay, by)),
                                         __builtin_fmaxf(__builtin_fminf(az, bz), 0.0f));
        const float tf = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(ax, bx), __builtin_fmaxf(ay, by)),
                                         bvh_min_raw(__builtin_fmaxf(az, bz), tlim));
        uint32_t next = w.w & 0xffffu;
        if (tn <= tf) {
            const uint32_t leaf = w.w >> 16;
            if (leaf == 0u) {
                next = i + 1u;
            } else {
                const uint32_t first = (leaf >> 2) - 1u, cnt = (leaf & 3u) + 1u;
                for (uint32_t k = 0; k < cnt; ++k) {
                    if (filter_sphere(lsph[first + k], kLeafSlot | (int32_t)(first + k), o, d, a, f))
                        tlim = bvh_min_raw(tlim, f.best_hi * inv_a_up);
                }
                q.tests += cnt;
                // t of the best sphere <= tlim, and tlim*tlim < dist2 proves it occludes
                if (shadow && tlim * tlim < dist2)
                    next = n_nodes;
            }
        }
        i = next;
    }
    if constexpr (kPriority)
        __builtin_amdgcn_s_setprio(0);

    Hit h;
    if (f.amb != 0u) {
        ++q.exact;
        q.tests += n_all;
        h = trace_exact(flat, n_all, o, d);
    } else {
        const int32_t best = f.best >= kLeafSlot ? lidx[f.best - kLeafSlot] : f.best;
        h.id = best;
        h.t = best >= 0 ? exact_t(flat((uint32_t)best), o, d, a) : FLT_MAX;
    }
    return h;
}

}  // namespace atx
