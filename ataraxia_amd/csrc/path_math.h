// path_math.h -- scalar arithmetic of the per-pixel path, shared by the HIP
// kernels and by the host-side camera code of libatx.
//
// Every function rounds exactly as the reference's GLM expressions do under
// IEEE-754 binary32 with no contraction (this file must be compiled with
// -ffp-contract=off and without fast-math; fused multiply-adds appear only
// where written as fmaf, in the sin/cos kernels).  GLM semantics restated
// (GLM is an unvendored dependency of the reference, see .gitmodules:1-4):
//   dot(a,b)     = (a.x*b.x + a.y*b.y) + a.z*b.z
//   normalize(v) = v * (1 / sqrt(dot(v,v)))
//   max(a,b)     = a < b ? b : a,  min(a,b) = b < a ? b : a
//   mix(x,y,a)   = x*(1-a) + y*a
//   mat4*vec4    = (m0*x + m1*y) + (m2*z + m3*w)
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ATX_HD __host__ __device__ __forceinline__
#else
#include <math.h>
#define ATX_HD inline
#endif

#if defined(__FAST_MATH__)
#error "path arithmetic must not be compiled with fast-math"
#endif

namespace atx {

struct f3 {
    float x, y, z;
};

ATX_HD f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
ATX_HD f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
ATX_HD f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
ATX_HD f3 mul(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
ATX_HD f3 scale(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
ATX_HD f3 divs(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
ATX_HD f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
ATX_HD float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }

#if defined(__HIP_DEVICE_COMPILE__)
ATX_HD float sqrt_cr(float x) { return __builtin_sqrtf(x); }   // correctly rounded (HIP default)
ATX_HD float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
ATX_HD float rint_(float x) { return __builtin_rintf(x); }
ATX_HD float fabs_(float x) { return __builtin_fabsf(x); }
#else
ATX_HD float sqrt_cr(float x) { return sqrtf(x); }
ATX_HD float fma_(float a, float b, float c) { return fmaf(a, b, c); }
ATX_HD float rint_(float x) { return rintf(x); }
ATX_HD float fabs_(float x) { return fabsf(x); }
#endif

ATX_HD float length(f3 a) { return sqrt_cr(dot(a, a)); }
ATX_HD f3 normalize(f3 a) { float inv = 1.0f / sqrt_cr(dot(a, a)); return scale(a, inv); }
ATX_HD f3 cross(f3 a, f3 b)
{
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
ATX_HD float gmax(float a, float b) { return a < b ? b : a; }
ATX_HD float gmin(float a, float b) { return b < a ? b : a; }
ATX_HD float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

constexpr float kPi = 3.14159274101257324219f;  // glm::pi<float>()

// ---- PCG hash (Core/include/Random.h:59-70) --------------------------------
ATX_HD uint32_t pcg_hash(uint32_t seed)
{
    uint32_t state = seed * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
// float(seed) / float(UINT32_MAX): float(UINT32_MAX) rounds to 2^32, so the
// quotient is float(seed) * 2^-32 exactly.
ATX_HD float pcg_float(uint32_t& seed)
{
    seed = pcg_hash(seed);
    return (float)seed * 0x1p-32f;
}

// ---- sin / cos: Cody-Waite reduction + Taylor polynomials ------------------
// x = k*pi/2 + r, pi/2 = C1 + C2 + C3 subtracted with one fma per part; sin
// to degree 9 and cos to degree 10 on [-pi/4, pi/4].  Accurate to about one
// ulp for |x| < 100; the path only evaluates x in [0, 2*pi].
ATX_HD void sincos_(float x, float& s, float& c)
{
    const float k = rint_(x * 0x1.45f306p-1f);
    float r = fma_(-k, 0x1.921fb6p+0f, x);
    r = fma_(-k, -0x1.777a5cp-25f, r);
    r = fma_(-k, -0x1.ee59dap-50f, r);
    const float r2 = r * r;
    float p = 0x1.71de3ap-19f;
    p = fma_(p, r2, -0x1.a01a02p-13f);
    p = fma_(p, r2, 0x1.111112p-7f);
    p = fma_(p, r2, -0x1.555556p-3f);
    const float r3 = r2 * r;
    const float ps = fma_(r3, p, r);
    float q = -0x1.27e4fcp-22f;
    q = fma_(q, r2, 0x1.a01a02p-16f);
    q = fma_(q, r2, -0x1.6c16c2p-10f);
    q = fma_(q, r2, 0x1.555556p-5f);
    q = fma_(q, r2, -0.5f);
    const float pc = fma_(q, r2, 1.0f);
    const int quad = ((int)k) & 3;
    const float a = (quad & 1) ? pc : ps;     // sin magnitude source
    const float b = (quad & 1) ? ps : pc;     // cos magnitude source
    s = (quad & 2) ? -a : a;
    c = ((quad + 1) & 2) ? -b : b;
}

ATX_HD float pow5(float x)
{
    const float x2 = x * x;
    const float x4 = x2 * x2;
    return x4 * x;
}

// ---- camera (Engine/src/Camera.cpp:134-195) ---------------------------------
// Column-major 4x4, m[col*4+row].
ATX_HD void mat_vec4(const float* m, float x, float y, float z, float w, float out[4])
{
    for (int r = 0; r < 4; ++r) {
        const float mul0 = m[0 * 4 + r] * x;
        const float mul1 = m[1 * 4 + r] * y;
        const float mul2 = m[2 * 4 + r] * z;
        const float mul3 = m[3 * 4 + r] * w;
        out[r] = (mul0 + mul1) + (mul2 + mul3);
    }
}

// Primary ray direction of pixel (x, y): coord = (x/W, y/H)*2 - 1 (no pixel
// centre offset, no jitter), unprojected and rotated into world space.
ATX_HD f3 camera_ray(const float* inv_proj, const float* inv_view, uint32_t x, uint32_t y,
                     uint32_t width, uint32_t height)
{
    float cx = (float)x / (float)width;
    float cy = (float)y / (float)height;
    cx = cx * 2.0f - 1.0f;
    cy = cy * 2.0f - 1.0f;
    float t[4];
    mat_vec4(inv_proj, cx, cy, 1.0f, 1.0f, t);
    const f3 t3 = mk(t[0] / t[3], t[1] / t[3], t[2] / t[3]);
    const f3 n = normalize(t3);
    float w[4];
    mat_vec4(inv_view, n.x, n.y, n.z, 0.0f, w);
    return normalize(mk(w[0], w[1], w[2]));
}

// ---- BRDF (Engine/src/BRDF.cu:9-70) ----------------------------------------
ATX_HD float distribution_ggx(float n_dot_h, float roughness)
{
    const float a = roughness * roughness;
    const float a2 = a * a;
    const float n_dot_h2 = n_dot_h * n_dot_h;
    float denom = n_dot_h2 * (a2 - 1.0f) + 1.0f;
    denom = kPi * denom * denom;
    return a2 / denom;
}

ATX_HD float geometry_schlick_ggx(float n_dot_v, float roughness)
{
    const float r = roughness + 1.0f;
    const float k = (r * r) / 8.0f;
    return n_dot_v / (n_dot_v * (1.0f - k) + k);
}

ATX_HD f3 cook_torrance(f3 albedo, f3 F0, float metallic, float roughness, f3 N, f3 V, f3 L)
{
    const f3 H = normalize(add(V, L));
    const float n_dot_l = gmax(dot(N, L), 0.0000001f);
    const float n_dot_v = gmax(dot(N, V), 0.0000001f);
    const float n_dot_h = gmax(dot(N, H), 0.0f);
    const float v_dot_h = gmax(dot(V, H), 0.0f);
    const float cos_theta = gclamp(v_dot_h, 0.0f, 1.0f);
    const float p5 = pow5(1.0f - cos_theta);
    const f3 one = mk(1.0f, 1.0f, 1.0f);
    const f3 F = add(F0, scale(sub(one, F0), p5));
    const float D = distribution_ggx(n_dot_h, roughness);
    const float G = geometry_schlick_ggx(n_dot_v, roughness) * geometry_schlick_ggx(n_dot_l, roughness);
    const f3 kD = scale(sub(one, F), 1.0f - metallic);
    const f3 num = scale(F, D * G);
    const float denom = 4.0f * n_dot_l * n_dot_v + 0.001f;
    const f3 specular = divs(num, denom);
    const f3 diffuse = divs(mul(sub(one, F), albedo), kPi);
    return scale(add(mul(kD, diffuse), specular), n_dot_l);
}

// ---- samplers (Engine/src/BRDF.cu:72-116) ------------------------------------
ATX_HD f3 to_world(f3 N, float hx, float hy, float hz)
{
    f3 T;
    if (fabs_(N.x) > fabs_(N.y))
        T = divs(mk(-N.z, 0.0f, N.x), sqrt_cr(N.x * N.x + N.z * N.z));
    else
        T = divs(mk(0.0f, -N.z, N.y), sqrt_cr(N.y * N.y + N.z * N.z));
    const f3 B = cross(N, T);
    return add(add(scale(T, hx), scale(B, hy)), scale(N, hz));
}

// The next direction of a path (Renderer.cu:381-384): the GGX half-vector
// when the material is metallic (BRDF.cu:95-116: cos theta =
// sqrt((1 - u1) / (1 + (a^2 - 1) u1)), a = roughness^2, sin theta =
// sqrt(1 - cos^2 theta)), else a cosine-weighted direction (BRDF.cu:72-93:
// radius sqrt(u1), height sqrt(1 - u1)).  Both take u1 and u2 from the stream
// in that order and place (radius cos phi, radius sin phi, height) with
// phi = 2 pi u2 in the same frame, so only radius and height depend on the
// branch: written this way a wave whose lanes differ in material runs the
// angle and the frame once.
ATX_HD f3 sample_next(f3 N, bool ggx, float roughness, uint32_t& seed)
{
    const float u1 = pcg_float(seed);
    const float u2 = pcg_float(seed);
    float radius, height;
    if (ggx) {
        const float a = roughness * roughness;
        const float cos_theta = sqrt_cr((1.0f - u1) / (1.0f + (a * a - 1.0f) * u1));
        radius = sqrt_cr(1.0f - cos_theta * cos_theta);
        height = cos_theta;
    } else {
        radius = sqrt_cr(u1);
        height = sqrt_cr(1.0f - u1);
    }
    const float phi = 2.0f * kPi * u2;
    float s, c;
    sincos_(phi, s, c);
    return to_world(N, radius * c, radius * s, height);
}

// ---- display (Engine/include/Renderer.h:70-78) -------------------------------
ATX_HD uint32_t to_u8(float v)
{
    const float s = v * 255.0f;
    return (s >= 0.0f) ? ((uint32_t)s & 0xFFu) : 0u;   // NaN -> 0
}
ATX_HD uint32_t rgba8(float r, float g, float b, float a)
{
    return (to_u8(a) << 24) | (to_u8(b) << 16) | (to_u8(g) << 8) | to_u8(r);
}

}  // namespace atx
