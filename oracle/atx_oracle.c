/*
 * atx_oracle.c -- CPU restatement of Ataraxia's per-pixel path tracer.
 *
 * TEST INFRASTRUCTURE ONLY (see atx_oracle.h).  Built with
 *   gcc -O3 -mfma -ffp-contract=off -fno-fast-math -fPIC -shared
 * so that every float expression below rounds exactly as written.
 *
 * Reference files restated (paths relative to the reference repository):
 *   Core/include/Random.h:59-70         PcgHash, PcgFloat
 *   Engine/src/Camera.cpp:134-195       projection/view matrices, ray directions
 *   Engine/src/Renderer.cu:150-170      kernelRender accumulation + display
 *   Engine/src/Renderer.cu:251-285      traceRay
 *   Engine/src/Renderer.cu:287-387      perPixel
 *   Engine/src/Renderer.cu:389-409      rayMiss / rayHit
 *   Engine/src/BRDF.cu:9-116            cookTorrance and helpers, samplers
 *   Engine/include/Renderer.h:70-78     vec4ToRGBA
 *   Engine/src/Renderer.cu:67-96        traverseSceneGraph (flatten)
 *   Engine/src/SceneNode.cpp:42-59      updateGlobalTransform
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   pcg_hash/pcg_float    independent Python-integer restatement of the
 *                         published PCG-RXS-M-XS hash + committed vectors
 *   sin/cos/pow5          max-ulp error against float64 libm over dense samples
 *   perspective/look_at   closed forms of the textbook RH_NO matrices (numpy f64)
 *   inverse               M * inverse(M) == I, and agreement with numpy.linalg.inv
 *   ray directions        unit length, centre pixel ~ forward, FOV edge angle
 *   trace_ray             head-on t = d - r, tangent rays, origin inside ignored,
 *                         lowest index wins exact ties, brute-force float64 argmin
 *   cook_torrance         F(1)=F0, D(1)=1/(pi a2), G(1,1)=1 corners, float64
 *                         closed form of the same formula for random inputs
 *   samplers              E[z]=2/3 and E[z^2]=1/2 (cosine); GGX cos(theta) CDF
 *   per_pixel             sky-only scene == (0.6,0.7,0.9); emissive sphere with one
 *                         bounce == emission; one diffuse sphere + one light with
 *                         one bounce == float64 evaluation of the path by hand;
 *                         occluded light == 0; no lights; zero bounces == 0;
 *                         multi-bounce paths (up to 6 bounces, diffuse and
 *                         metallic, two lights, sky) against a separate float64
 *                         restatement with Python-integer PCG (tests/_pathref.py):
 *                         per-bounce hit and shadow ids, seeds after `seed += i`,
 *                         light indices and roulette draws exact, p, directions
 *                         and radiance to rounding; closed forms: albedo 1 under
 *                         the sky == sky exactly, albedo 0.5 == sky*0.5/p or 0 by
 *                         the hash-chain draw, first continuation cos(theta) from
 *                         u1 for GGX and cosine
 *   render_frame          accumulation mean, RGBA8 truncation table, Kahan vs
 *                         naive vs float64 sums, golden images in tests/golden
 *   flatten               sample-scene-data/scene.json known answers; closed
 *                         forms of quaternion order ([x,y,z,w] in the file),
 *                         90-degree rotations, T*R*S order, parent*local order
 *                         and the mean-column-length radius scale
 * No oracle function is left without one of these checks.
 */
#include "atx_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#if defined(__FP_FAST_FMAF) || defined(__FAST_MATH__)
#  if defined(__FAST_MATH__)
#    error "the oracle must not be built with fast-math"
#  endif
#endif

/* ------------------------------------------------------------------------- */
/* GLM semantics (a13)                                                        */
/* ------------------------------------------------------------------------- */
typedef struct { float x, y, z; } v3;

static v3 v3make(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static v3 v3load(const float* p) { return v3make(p[0], p[1], p[2]); }
static v3 v3add(v3 a, v3 b) { return v3make(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 v3sub(v3 a, v3 b) { return v3make(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 v3mul(v3 a, v3 b) { return v3make(a.x * b.x, a.y * b.y, a.z * b.z); }
static v3 v3scale(v3 a, float s) { return v3make(a.x * s, a.y * s, a.z * s); }
static v3 v3divs(v3 a, float s) { return v3make(a.x / s, a.y / s, a.z / s); }
static v3 v3neg(v3 a) { return v3make(-a.x, -a.y, -a.z); }
static float v3dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static float v3len(v3 a) { return sqrtf(v3dot(a, a)); }
static v3 v3norm(v3 a) { float inv = 1.0f / sqrtf(v3dot(a, a)); return v3scale(a, inv); }
static v3 v3cross(v3 a, v3 b)
{
    return v3make(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static float gmax(float a, float b) { return a < b ? b : a; }
static float gmin(float a, float b) { return b < a ? b : a; }
static float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }

#define ORA_PI_F 3.14159274101257324219f /* glm::pi<float>() */

/* ------------------------------------------------------------------------- */
/* a1/a2: PCG (Core/include/Random.h:59-70)                                   */
/* ------------------------------------------------------------------------- */
uint32_t ora_pcg_hash(uint32_t seed)
{
    uint32_t state = seed * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

float ora_pcg_float(uint32_t* seed)
{
    *seed = ora_pcg_hash(*seed);
    return (float)(*seed) / (float)UINT32_MAX;
}

/* ------------------------------------------------------------------------- */
/* sin / cos / pow5: the project's shared implementations                     */
/* ------------------------------------------------------------------------- */
/* Cody & Waite (1980) reduction x = k*pi/2 + r with pi/2 = C1 + C2 + C3 and a
 * fused multiply-add per part, then Taylor polynomials on [-pi/4, pi/4]:
 * sin to degree 9 (truncation < 2e-9 relative), cos to degree 10 (< 1e-10).  */
static const float CW_C1 = 0x1.921fb6p+0f;
static const float CW_C2 = -0x1.777a5cp-25f;
static const float CW_C3 = -0x1.ee59dap-50f;
static const float TWO_OVER_PI = 0x1.45f306p-1f;

static float poly_sin(float r)
{
    float r2 = r * r;
    float p = 0x1.71de3ap-19f;              /*  1/9! */
    p = fmaf(p, r2, -0x1.a01a02p-13f);      /* -1/7! */
    p = fmaf(p, r2, 0x1.111112p-7f);        /*  1/5! */
    p = fmaf(p, r2, -0x1.555556p-3f);       /* -1/3! */
    float r3 = r2 * r;
    return fmaf(r3, p, r);
}

static float poly_cos(float r)
{
    float r2 = r * r;
    float q = -0x1.27e4fcp-22f;             /* -1/10! */
    q = fmaf(q, r2, 0x1.a01a02p-16f);       /*  1/8!  */
    q = fmaf(q, r2, -0x1.6c16c2p-10f);      /* -1/6!  */
    q = fmaf(q, r2, 0x1.555556p-5f);        /*  1/4!  */
    q = fmaf(q, r2, -0.5f);                 /* -1/2!  */
    return fmaf(q, r2, 1.0f);
}

static void ora_sincos(float x, float* s, float* c)
{
    float k = rintf(x * TWO_OVER_PI);
    float r = fmaf(-k, CW_C1, x);
    r = fmaf(-k, CW_C2, r);
    r = fmaf(-k, CW_C3, r);
    float ps = poly_sin(r), pc = poly_cos(r);
    int q = ((int)k) & 3;
    switch (q) {
    case 0: *s = ps; *c = pc; break;
    case 1: *s = pc; *c = -ps; break;
    case 2: *s = -ps; *c = -pc; break;
    default: *s = -pc; *c = ps; break;
    }
}

float ora_sin(float x) { float s, c; ora_sincos(x, &s, &c); return s; }
float ora_cos(float x) { float s, c; ora_sincos(x, &s, &c); return c; }

float ora_pow5(float x)
{
    float x2 = x * x;
    float x4 = x2 * x2;
    return x4 * x;
}

/* ------------------------------------------------------------------------- */
/* a4: camera matrices and ray directions (Engine/src/Camera.cpp:134-195)     */
/* ------------------------------------------------------------------------- */
static void mat_identity(float m[16])
{
    memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.0f;
}

/* glm::perspective, right-handed, clip z in [-1, 1]. */
void ora_perspective(float fov_deg, float aspect, float znear, float zfar, float out[16])
{
    float fovy = fov_deg * 0x1.1df46ap-6f; /* glm::radians */
    float tan_half = tanf(fovy / 2.0f);
    memset(out, 0, 16 * sizeof(float));
    out[0 * 4 + 0] = 1.0f / (aspect * tan_half);
    out[1 * 4 + 1] = 1.0f / tan_half;
    out[2 * 4 + 2] = -(zfar + znear) / (zfar - znear);
    out[2 * 4 + 3] = -1.0f;
    out[3 * 4 + 2] = -(2.0f * zfar * znear) / (zfar - znear);
}

/* glm::lookAt, right-handed. */
void ora_look_at(const float eye_[3], const float center_[3], const float up_[3], float out[16])
{
    v3 eye = v3load(eye_), center = v3load(center_), up = v3load(up_);
    v3 f = v3norm(v3sub(center, eye));
    v3 s = v3norm(v3cross(f, up));
    v3 u = v3cross(s, f);
    mat_identity(out);
    out[0 * 4 + 0] = s.x; out[1 * 4 + 0] = s.y; out[2 * 4 + 0] = s.z;
    out[0 * 4 + 1] = u.x; out[1 * 4 + 1] = u.y; out[2 * 4 + 1] = u.z;
    out[0 * 4 + 2] = -f.x; out[1 * 4 + 2] = -f.y; out[2 * 4 + 2] = -f.z;
    out[3 * 4 + 0] = -v3dot(s, eye);
    out[3 * 4 + 1] = -v3dot(u, eye);
    out[3 * 4 + 2] = v3dot(f, eye);
}

/* glm::inverse for mat4: cofactor expansion with 2x2 sub-determinants
 * ("Coef"), signed adjugate columns, determinant from the first row.       */
void ora_inverse(const float mm[16], float out[16])
{
#define M(c, r) mm[(c) * 4 + (r)]
    float c00 = M(2,2) * M(3,3) - M(3,2) * M(2,3);
    float c02 = M(1,2) * M(3,3) - M(3,2) * M(1,3);
    float c03 = M(1,2) * M(2,3) - M(2,2) * M(1,3);
    float c04 = M(2,1) * M(3,3) - M(3,1) * M(2,3);
    float c06 = M(1,1) * M(3,3) - M(3,1) * M(1,3);
    float c07 = M(1,1) * M(2,3) - M(2,1) * M(1,3);
    float c08 = M(2,1) * M(3,2) - M(3,1) * M(2,2);
    float c10 = M(1,1) * M(3,2) - M(3,1) * M(1,2);
    float c11 = M(1,1) * M(2,2) - M(2,1) * M(1,2);
    float c12 = M(2,0) * M(3,3) - M(3,0) * M(2,3);
    float c14 = M(1,0) * M(3,3) - M(3,0) * M(1,3);
    float c15 = M(1,0) * M(2,3) - M(2,0) * M(1,3);
    float c16 = M(2,0) * M(3,2) - M(3,0) * M(2,2);
    float c18 = M(1,0) * M(3,2) - M(3,0) * M(1,2);
    float c19 = M(1,0) * M(2,2) - M(2,0) * M(1,2);
    float c20 = M(2,0) * M(3,1) - M(3,0) * M(2,1);
    float c22 = M(1,0) * M(3,1) - M(3,0) * M(1,1);
    float c23 = M(1,0) * M(2,1) - M(2,0) * M(1,1);

    float fac0[4] = {c00, c00, c02, c03};
    float fac1[4] = {c04, c04, c06, c07};
    float fac2[4] = {c08, c08, c10, c11};
    float fac3[4] = {c12, c12, c14, c15};
    float fac4[4] = {c16, c16, c18, c19};
    float fac5[4] = {c20, c20, c22, c23};
    float vec0[4] = {M(1,0), M(0,0), M(0,0), M(0,0)};
    float vec1[4] = {M(1,1), M(0,1), M(0,1), M(0,1)};
    float vec2[4] = {M(1,2), M(0,2), M(0,2), M(0,2)};
    float vec3[4] = {M(1,3), M(0,3), M(0,3), M(0,3)};
    float inv[4][4];
    static const float signA[4] = {+1.0f, -1.0f, +1.0f, -1.0f};
    static const float signB[4] = {-1.0f, +1.0f, -1.0f, +1.0f};
    for (int i = 0; i < 4; ++i) {
        float i0 = (vec1[i] * fac0[i] - vec2[i] * fac1[i]) + vec3[i] * fac2[i];
        float i1 = (vec0[i] * fac0[i] - vec2[i] * fac3[i]) + vec3[i] * fac4[i];
        float i2 = (vec0[i] * fac1[i] - vec1[i] * fac3[i]) + vec3[i] * fac5[i];
        float i3 = (vec0[i] * fac2[i] - vec1[i] * fac4[i]) + vec2[i] * fac5[i];
        inv[0][i] = i0 * signA[i];
        inv[1][i] = i1 * signB[i];
        inv[2][i] = i2 * signA[i];
        inv[3][i] = i3 * signB[i];
    }
    float row0[4] = {inv[0][0], inv[1][0], inv[2][0], inv[3][0]};
    float dot0[4] = {M(0,0) * row0[0], M(0,1) * row0[1], M(0,2) * row0[2], M(0,3) * row0[3]};
    float dot1 = (dot0[0] + dot0[1]) + (dot0[2] + dot0[3]);
    float one_over_det = 1.0f / dot1;
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r)
            out[c * 4 + r] = inv[c][r] * one_over_det;
#undef M
}

/* mat4 * vec4 = (m0*x + m1*y) + (m2*z + m3*w) */
static void mat_vec4(const float m[16], const float v[4], float out[4])
{
    for (int r = 0; r < 4; ++r) {
        float mul0 = m[0 * 4 + r] * v[0];
        float mul1 = m[1 * 4 + r] * v[1];
        float mul2 = m[2 * 4 + r] * v[2];
        float mul3 = m[3 * 4 + r] * v[3];
        out[r] = (mul0 + mul1) + (mul2 + mul3);
    }
}

void ora_camera_matrices(const ora_camera* cam, float inv_proj[16], float inv_view[16])
{
    float proj[16], view[16];
    float aspect = (float)cam->width / (float)cam->height;
    ora_perspective(cam->fov_deg, aspect, cam->near_clip, cam->far_clip, proj);
    ora_inverse(proj, inv_proj);
    v3 pos = v3load(cam->position), dir = v3load(cam->direction);
    v3 center = v3add(pos, dir);
    float c[3] = {center.x, center.y, center.z};
    float up[3] = {0.0f, 1.0f, 0.0f};
    ora_look_at(cam->position, c, up, view);
    ora_inverse(view, inv_view);
}

void ora_ray_direction(const float inv_proj[16], const float inv_view[16],
                       uint32_t x, uint32_t y, uint32_t width, uint32_t height, float out[3])
{
    float cx = (float)x / (float)width;
    float cy = (float)y / (float)height;
    cx = cx * 2.0f - 1.0f;
    cy = cy * 2.0f - 1.0f;
    float in4[4] = {cx, cy, 1.0f, 1.0f}, target[4];
    mat_vec4(inv_proj, in4, target);
    v3 t3 = v3make(target[0] / target[3], target[1] / target[3], target[2] / target[3]);
    v3 n = v3norm(t3);
    float d4[4] = {n.x, n.y, n.z, 0.0f}, w[4];
    mat_vec4(inv_view, d4, w);
    v3 d = v3norm(v3make(w[0], w[1], w[2]));
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

void ora_ray_directions(const ora_camera* cam, float* dirs)
{
    float ip[16], iv[16];
    ora_camera_matrices(cam, ip, iv);
    for (uint32_t y = 0; y < cam->height; ++y)
        for (uint32_t x = 0; x < cam->width; ++x)
            ora_ray_direction(ip, iv, x, y, cam->width, cam->height,
                              dirs + 3 * ((size_t)x + (size_t)y * cam->width));
}

/* ------------------------------------------------------------------------- */
/* a5/a6: traceRay, rayHit (Engine/src/Renderer.cu:251-285, 396-409)          */
/* ------------------------------------------------------------------------- */
int32_t ora_trace_ray(const float origin[3], const float dir[3],
                      const ora_sphere* spheres, uint64_t n, float* t_out)
{
    v3 o = v3load(origin), d = v3load(dir);
    int32_t closest = -1;
    float tmin = FLT_MAX;
    for (uint64_t i = 0; i < n; ++i) {
        v3 center = v3load(spheres[i].center);
        float radius = spheres[i].radius;
        v3 oc = v3sub(o, center);
        float a = v3dot(d, d);
        float b = 2.0f * v3dot(oc, d);
        float c = v3dot(oc, oc) - radius * radius;
        float disc = b * b - 4.0f * a * c;
        if (disc < 0.0f)
            continue;
        float t0 = (-b - sqrtf(disc)) / (2.0f * a);
        float t1 = (-b + sqrtf(disc)) / (2.0f * a);
        float t = t0 < t1 ? t0 : t1;
        if (t > 0.0f && t < tmin) {
            tmin = t;
            closest = (int32_t)i;
        }
    }
    if (t_out)
        *t_out = tmin;
    return closest;
}

typedef struct { float t; v3 pos; v3 normal; int32_t id; } hit_record;

static hit_record trace(v3 o, v3 d, const ora_scene* sc)
{
    hit_record h;
    float o3[3] = {o.x, o.y, o.z}, d3[3] = {d.x, d.y, d.z};
    float tmin;
    int32_t idx = ora_trace_ray(o3, d3, sc->spheres, sc->num_spheres, &tmin);
    if (idx < 0) { /* rayMiss */
        h.t = -1.0f;
        h.id = -1;
        h.pos = v3make(0.0f, 0.0f, 0.0f);
        h.normal = h.pos;
        return h;
    }
    v3 center = v3load(sc->spheres[idx].center);
    v3 oc = v3sub(o, center);
    h.t = tmin;
    h.id = idx;
    h.pos = v3add(oc, v3scale(d, tmin));
    h.normal = v3norm(h.pos);
    h.pos = v3add(h.pos, center);
    return h;
}

/* ------------------------------------------------------------------------- */
/* a8: Cook-Torrance (Engine/src/BRDF.cu:9-70)                                */
/* ------------------------------------------------------------------------- */
static float distribution_ggx(float n_dot_h, float roughness)
{
    float a = roughness * roughness;
    float a2 = a * a;
    float n_dot_h2 = n_dot_h * n_dot_h;
    float denom = n_dot_h2 * (a2 - 1.0f) + 1.0f;
    denom = ORA_PI_F * denom * denom;
    return a2 / denom;
}

static float geometry_schlick_ggx(float n_dot_v, float roughness)
{
    float r = roughness + 1.0f;
    float k = (r * r) / 8.0f;
    return n_dot_v / (n_dot_v * (1.0f - k) + k);
}

void ora_cook_torrance(const float albedo_[3], const float F0_[3], float metallic, float roughness,
                       const float N_[3], const float V_[3], const float L_[3], float out[3])
{
    v3 albedo = v3load(albedo_), F0 = v3load(F0_), N = v3load(N_), V = v3load(V_), L = v3load(L_);
    v3 H = v3norm(v3add(V, L));
    float n_dot_l = gmax(v3dot(N, L), 0.0000001f);
    float n_dot_v = gmax(v3dot(N, V), 0.0000001f);
    float n_dot_h = gmax(v3dot(N, H), 0.0f);
    float v_dot_h = gmax(v3dot(V, H), 0.0f);

    /* fresnelSchlick (BRDF.cu:36-40) */
    float cos_theta = gclamp(v_dot_h, 0.0f, 1.0f);
    float p5 = ora_pow5(1.0f - cos_theta);
    v3 one = v3make(1.0f, 1.0f, 1.0f);
    v3 F = v3add(F0, v3scale(v3sub(one, F0), p5));

    float D = distribution_ggx(n_dot_h, roughness);
    float G = geometry_schlick_ggx(n_dot_v, roughness) * geometry_schlick_ggx(n_dot_l, roughness);

    v3 kD = v3sub(one, F);
    kD = v3scale(kD, 1.0f - metallic);
    v3 num = v3scale(F, D * G);
    float denom = 4.0f * n_dot_l * n_dot_v + 0.001f;
    v3 specular = v3divs(num, denom);
    v3 diffuse = v3divs(v3mul(v3sub(one, F), albedo), ORA_PI_F);
    v3 r = v3scale(v3add(v3mul(kD, diffuse), specular), n_dot_l);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

/* ------------------------------------------------------------------------- */
/* a9: samplers (Engine/src/BRDF.cu:72-116)                                   */
/* ------------------------------------------------------------------------- */
static void tangent_frame(v3 N, v3* T, v3* B)
{
    if (fabsf(N.x) > fabsf(N.y))
        *T = v3divs(v3make(-N.z, 0.0f, N.x), sqrtf(N.x * N.x + N.z * N.z));
    else
        *T = v3divs(v3make(0.0f, -N.z, N.y), sqrtf(N.y * N.y + N.z * N.z));
    *B = v3cross(N, *T);
}

void ora_sample_cosine(const float N_[3], uint32_t* seed, float out[3])
{
    v3 N = v3load(N_);
    float u1 = ora_pcg_float(seed);
    float u2 = ora_pcg_float(seed);
    float r = sqrtf(u1);
    float theta = 2.0f * ORA_PI_F * u2;
    float x = r * ora_cos(theta);
    float y = r * ora_sin(theta);
    float z = sqrtf(1.0f - u1);
    v3 T, B;
    tangent_frame(N, &T, &B);
    v3 d = v3add(v3add(v3scale(T, x), v3scale(B, y)), v3scale(N, z));
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

void ora_sample_ggx(const float N_[3], float roughness, uint32_t* seed, float out[3])
{
    v3 N = v3load(N_);
    float u1 = ora_pcg_float(seed);
    float u2 = ora_pcg_float(seed);
    float a = roughness * roughness;
    float cos_theta = sqrtf((1.0f - u1) / (1.0f + (a * a - 1.0f) * u1));
    float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    float phi = 2.0f * ORA_PI_F * u2;
    v3 H = v3make(sin_theta * ora_cos(phi), sin_theta * ora_sin(phi), cos_theta);
    v3 T, B;
    tangent_frame(N, &T, &B);
    v3 d = v3add(v3add(v3scale(T, H.x), v3scale(B, H.y)), v3scale(N, H.z));
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}

/* ------------------------------------------------------------------------- */
/* a3/a7: perPixel (Engine/src/Renderer.cu:287-387)                           */
/* ------------------------------------------------------------------------- */
void ora_per_pixel(uint32_t x, uint32_t y, uint32_t width, const float dir[3],
                   const float cam_pos[3], const ora_scene* sc, uint32_t frame_index,
                   const ora_settings* settings, float out_rgba[4], ora_path_debug* dbg)
{
    v3 ro = v3load(cam_pos), rd = v3load(dir);
    v3 color = v3make(0.0f, 0.0f, 0.0f);
    v3 throughput = v3make(1.0f, 1.0f, 1.0f);
    uint32_t seed = x + y * width;
    seed *= frame_index;
    int bounces = settings->max_bounces;
    if (dbg) {
        dbg->rays = 0;
        dbg->bounces = 0;
        for (int i = 0; i < bounces; ++i) {
            if (dbg->hit_ids) dbg->hit_ids[i] = -2;
            if (dbg->shadow_ids) dbg->shadow_ids[i] = -2;
            if (dbg->seeds) dbg->seeds[i] = 0u;
            if (dbg->light_ids) dbg->light_ids[i] = -1;
            if (dbg->rr) dbg->rr[2 * i] = dbg->rr[2 * i + 1] = NAN;
            if (dbg->dirs) dbg->dirs[3 * i] = dbg->dirs[3 * i + 1] = dbg->dirs[3 * i + 2] = NAN;
        }
    }
    for (int i = 0; i < bounces; i++) {
        seed += (uint32_t)i;
        hit_record ht = trace(ro, rd, sc);
        if (dbg) {
            dbg->rays += 1;
            dbg->bounces = i + 1;
            if (dbg->hit_ids) dbg->hit_ids[i] = ht.id;
            if (dbg->seeds) dbg->seeds[i] = seed;
            if (dbg->dirs) { dbg->dirs[3 * i] = rd.x; dbg->dirs[3 * i + 1] = rd.y; dbg->dirs[3 * i + 2] = rd.z; }
        }
        if (ht.t < 0.0f) {
            if (settings->sky_light) {
                v3 miss = v3make(0.6f, 0.7f, 0.9f);
                color = v3add(color, v3mul(miss, throughput));
            }
            break;
        }
        const ora_material* mat = &sc->materials[sc->spheres[ht.id].material];
        v3 albedo = v3load(mat->albedo);
        if (mat->emission_intensity > 0.0f) {
            v3 emission = v3scale(v3load(mat->emission_color), mat->emission_intensity);
            color = v3add(color, v3mul(emission, throughput));
        }
        /* glm::mix(F0, albedo, metallic) */
        v3 base_refl = v3add(v3scale(v3load(mat->F0), 1.0f - mat->metallic), v3scale(albedo, mat->metallic));

        if (sc->num_lights > 0) {
            uint32_t light_index = (uint32_t)((uint64_t)ora_pcg_hash(seed) % sc->num_lights);
            if (dbg && dbg->light_ids) dbg->light_ids[i] = (int32_t)light_index;
            const ora_light* light = &sc->lights[light_index];
            v3 L = v3sub(v3load(light->position), ht.pos);
            float dist2 = v3dot(L, L);
            L = v3norm(L);
            v3 so = v3add(ht.pos, v3scale(ht.normal, 0.0001f));
            hit_record sh = trace(so, L, sc);
            if (dbg) {
                dbg->rays += 1;
                if (dbg->shadow_ids) dbg->shadow_ids[i] = sh.id;
            }
            if (sh.t > 0.0f && sh.t * sh.t < dist2) {
                /* occluded */
            } else {
                v3 V = v3neg(rd);
                float f[3], a3[3] = {albedo.x, albedo.y, albedo.z};
                float F0p[3] = {base_refl.x, base_refl.y, base_refl.z};
                float n3[3] = {ht.normal.x, ht.normal.y, ht.normal.z};
                float v3a[3] = {V.x, V.y, V.z}, l3[3] = {L.x, L.y, L.z};
                ora_cook_torrance(a3, F0p, mat->metallic, mat->roughness, n3, v3a, l3, f);
                v3 spec = v3load(f);
                v3 emission = v3scale(v3load(light->color), light->intensity);
                float pdf = 1.0f;
                color = v3add(color, v3divs(v3mul(v3mul(emission, spec), throughput), pdf));
            }
        }

        throughput = v3mul(throughput, albedo);
        ro = v3add(ht.pos, v3scale(ht.normal, 0.0001f));

        float p = gmax(0.1f, gmin(1.0f, v3len(throughput)));
        float u = ora_pcg_float(&seed);
        if (dbg && dbg->rr) { dbg->rr[2 * i] = p; dbg->rr[2 * i + 1] = u; }
        if (u > p)
            break;
        throughput = v3divs(throughput, p);

        float n3[3] = {ht.normal.x, ht.normal.y, ht.normal.z}, nd[3];
        if (mat->metallic > 0.0f)
            ora_sample_ggx(n3, mat->roughness, &seed, nd);
        else
            ora_sample_cosine(n3, &seed, nd);
        rd = v3load(nd);
    }
    out_rgba[0] = color.x;
    out_rgba[1] = color.y;
    out_rgba[2] = color.z;
    out_rgba[3] = 1.0f;
}

/* ------------------------------------------------------------------------- */
/* a10: accumulation and display (Renderer.cu:150-170, Renderer.h:70-78)      */
/* ------------------------------------------------------------------------- */
static uint32_t to_u8(float v)
{
    float s = v * 255.0f;
    if (!(s >= 0.0f)) /* NaN; clamp already maps the rest into [0, 255] */
        return 0u;
    return (uint32_t)s & 0xFFu;
}

uint32_t ora_rgba8(const float c[4])
{
    uint32_t r = to_u8(c[0]), g = to_u8(c[1]), b = to_u8(c[2]), a = to_u8(c[3]);
    return (a << 24) | (b << 16) | (g << 8) | r;
}

void ora_render_frame(const ora_camera* cam, const ora_scene* sc, const ora_settings* settings,
                      uint32_t frame_index, float* accum, float* comp, uint32_t* image,
                      float* radiance, uint32_t row0, uint32_t row1, uint64_t* rays)
{
    float ip[16], iv[16];
    ora_camera_matrices(cam, ip, iv);
    uint32_t W = cam->width;
    uint64_t nrays = 0;
    for (uint32_t y = row0; y < row1; ++y) {
        for (uint32_t x = 0; x < W; ++x) {
            float d[3], c[4];
            ora_ray_direction(ip, iv, x, y, W, cam->height, d);
            ora_path_debug dbg = {0};
            ora_per_pixel(x, y, W, d, cam->position, sc, frame_index, settings, c, rays ? &dbg : 0);
            nrays += dbg.rays;
            size_t px = (size_t)x + (size_t)y * W;
            float* acc = accum + 4 * px;
            if (radiance)
                memcpy(radiance + 4 * px, c, sizeof(c));
            if (comp) {
                float* cp = comp + 4 * px;
                for (int k = 0; k < 4; ++k) {
                    float yk = c[k] - cp[k];
                    float t = acc[k] + yk;
                    cp[k] = (t - acc[k]) - yk;
                    acc[k] = t;
                }
            } else {
                for (int k = 0; k < 4; ++k)
                    acc[k] = acc[k] + c[k];
            }
            float fin[4];
            for (int k = 0; k < 4; ++k)
                fin[k] = gclamp(acc[k] / (float)frame_index, 0.0f, 1.0f);
            if (image)
                image[px] = ora_rgba8(fin);
        }
    }
    if (rays)
        *rays = nrays;
}

void ora_render_frame_debug(const ora_camera* cam, const ora_scene* sc, const ora_settings* settings,
                            uint32_t frame_index, float* radiance, int32_t* hit_ids, int32_t* shadow_ids,
                            uint32_t max_bounces, uint32_t* rays, uint32_t row0, uint32_t row1)
{
    float ip[16], iv[16];
    ora_camera_matrices(cam, ip, iv);
    uint32_t W = cam->width;
    int nb = settings->max_bounces > 0 ? settings->max_bounces : 1;
    int32_t hbuf[64], sbuf[64];
    int32_t* hb = nb <= 64 ? hbuf : (int32_t*)malloc(sizeof(int32_t) * (size_t)nb);
    int32_t* sb = nb <= 64 ? sbuf : (int32_t*)malloc(sizeof(int32_t) * (size_t)nb);
    for (uint32_t y = row0; y < row1; ++y) {
        for (uint32_t x = 0; x < W; ++x) {
            float d[3], c[4];
            size_t px = (size_t)x + (size_t)y * W;
            ora_ray_direction(ip, iv, x, y, W, cam->height, d);
            ora_path_debug dbg = {0};
            dbg.hit_ids = hb;
            dbg.shadow_ids = sb;
            ora_per_pixel(x, y, W, d, cam->position, sc, frame_index, settings, c, &dbg);
            if (radiance) memcpy(radiance + 4 * px, c, sizeof(c));
            for (uint32_t b = 0; b < max_bounces; ++b) {
                int32_t hv = (int)b < settings->max_bounces ? hb[b] : -2;
                int32_t sv = (int)b < settings->max_bounces ? sb[b] : -2;
                if (hit_ids) hit_ids[px * max_bounces + b] = hv;
                if (shadow_ids) shadow_ids[px * max_bounces + b] = sv;
            }
            if (rays) rays[px] = (uint32_t)dbg.rays;
        }
    }
    if (hb != hbuf) free(hb);
    if (sb != sbuf) free(sb);
}

typedef struct {
    const ora_camera* cam; const ora_scene* sc; const ora_settings* settings;
    uint32_t frame; float* accum; float* comp; uint32_t* image;
    volatile uint32_t* next_row; uint64_t rays;
} mt_job;

static void* mt_worker(void* arg)
{
    mt_job* j = (mt_job*)arg;
    for (;;) {
        uint32_t row = __sync_fetch_and_add(j->next_row, 4u);
        if (row >= j->cam->height)
            break;
        uint32_t end = row + 4u < j->cam->height ? row + 4u : j->cam->height;
        uint64_t r = 0;
        ora_render_frame(j->cam, j->sc, j->settings, j->frame, j->accum, j->comp, j->image, 0,
                         row, end, &r);
        j->rays += r;
    }
    return 0;
}

int ora_render_frames_mt(const ora_camera* cam, const ora_scene* sc, const ora_settings* settings,
                         uint32_t first_frame, uint32_t frames, float* accum, float* comp,
                         uint32_t* image, int threads, uint64_t* rays)
{
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    size_t n = (size_t)cam->width * cam->height * 4;
    uint64_t total = 0;
    for (uint32_t f = 0; f < frames; ++f) {
        uint32_t frame = first_frame + f;
        if (frame == 1) {
            memset(accum, 0, n * sizeof(float));
            if (comp) memset(comp, 0, n * sizeof(float));
        }
        volatile uint32_t next_row = 0;
        pthread_t tid[256];
        mt_job jobs[256];
        for (int t = 0; t < threads; ++t) {
            mt_job jb = {cam, sc, settings, frame, accum, comp, image, &next_row, 0};
            jobs[t] = jb;
            if (pthread_create(&tid[t], 0, mt_worker, &jobs[t]) != 0)
                return -1;
        }
        for (int t = 0; t < threads; ++t) {
            pthread_join(tid[t], 0);
            total += jobs[t].rays;
        }
    }
    if (rays) *rays = total;
    return 0;
}

typedef struct {
    const ora_camera* cam; const ora_scene* sc; const ora_settings* settings;
    const float* ip; const float* iv;
    uint32_t first_frame, frames; const uint32_t* pixels; uint64_t n; int kahan;
    float* accum; uint32_t* image; volatile uint64_t* next;
    double* accum64;   /* float64 sums instead (accum and image unused) */
} px_job;

static void* px_worker(void* arg)
{
    px_job* j = (px_job*)arg;
    uint32_t W = j->cam->width;
    for (;;) {
        uint64_t i = __sync_fetch_and_add(j->next, 1u);
        if (i >= j->n)
            break;
        uint32_t x = j->pixels[i] % W, y = j->pixels[i] / W;
        float d[3], acc[4] = {0, 0, 0, 0}, cp[4] = {0, 0, 0, 0};
        double a64[4] = {0, 0, 0, 0};
        ora_ray_direction(j->ip, j->iv, x, y, W, j->cam->height, d);
        /* Renderer.cu:245-248: the index advances with accumulation, else stays 1 */
        uint32_t fi = j->first_frame;
        for (uint32_t f = 0; f < j->frames; ++f) {
            float c[4];
            ora_per_pixel(x, y, W, d, j->cam->position, j->sc, fi, j->settings, c, 0);
            if (fi == 1u)
                for (int k = 0; k < 4; ++k) { acc[k] = cp[k] = 0.0f; a64[k] = 0.0; }
            for (int k = 0; k < 4; ++k) {
                if (j->accum64) {
                    a64[k] += (double)c[k];
                } else if (j->kahan) {
                    float yk = c[k] - cp[k];
                    float t = acc[k] + yk;
                    cp[k] = (t - acc[k]) - yk;
                    acc[k] = t;
                } else {
                    acc[k] = acc[k] + c[k];
                }
            }
            if (f + 1 < j->frames)
                fi = j->settings->accumulation ? fi + 1u : 1u;
        }
        if (j->accum64) {
            memcpy(j->accum64 + 4 * i, a64, sizeof(a64));
            continue;
        }
        memcpy(j->accum + 4 * i, acc, sizeof(acc));
        if (j->image) {
            float fin[4];
            for (int k = 0; k < 4; ++k)
                fin[k] = gclamp(acc[k] / (float)fi, 0.0f, 1.0f);
            j->image[i] = ora_rgba8(fin);
        }
    }
    return 0;
}

static int accumulate_pixels(const ora_camera* cam, const ora_scene* sc, const ora_settings* settings,
                             uint32_t first_frame, uint32_t frames, const uint32_t* pixels, uint64_t n,
                             int kahan, float* accum, uint32_t* image, double* accum64, int threads)
{
    float ip[16], iv[16];
    ora_camera_matrices(cam, ip, iv);
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    volatile uint64_t next = 0;
    pthread_t tid[256];
    px_job job = {cam, sc, settings, ip, iv, first_frame, frames, pixels, n, kahan, accum, image, &next, accum64};
    for (int t = 0; t < threads; ++t)
        if (pthread_create(&tid[t], 0, px_worker, &job) != 0) {
            for (int u = 0; u < t; ++u) pthread_join(tid[u], 0);
            return -1;
        }
    for (int t = 0; t < threads; ++t)
        pthread_join(tid[t], 0);
    return 0;
}

int ora_accumulate_pixels(const ora_camera* cam, const ora_scene* sc, const ora_settings* settings,
                          uint32_t first_frame, uint32_t frames, const uint32_t* pixels, uint64_t n,
                          int kahan, float* accum, uint32_t* image, int threads)
{
    return accumulate_pixels(cam, sc, settings, first_frame, frames, pixels, n, kahan, accum, image, 0, threads);
}

int ora_accumulate_pixels_f64(const ora_camera* cam, const ora_scene* sc, const ora_settings* settings,
                              uint32_t first_frame, uint32_t frames, const uint32_t* pixels, uint64_t n,
                              double* accum, int threads)
{
    return accumulate_pixels(cam, sc, settings, first_frame, frames, pixels, n, 0, 0, 0, accum, threads);
}

/* ------------------------------------------------------------------------- */
/* a11: scene-graph flattening (Renderer.cu:67-96, SceneNode.cpp:42-59)       */
/* ------------------------------------------------------------------------- */
void ora_mat4_mul(const float A[16], const float B[16], float out[16])
{
    float r[16];
    for (int c = 0; c < 4; ++c)
        for (int row = 0; row < 4; ++row)
            r[c * 4 + row] = ((A[0 * 4 + row] * B[c * 4 + 0] + A[1 * 4 + row] * B[c * 4 + 1])
                              + A[2 * 4 + row] * B[c * 4 + 2]) + A[3 * 4 + row] * B[c * 4 + 3];
    memcpy(out, r, sizeof(r));
}

void ora_node_local_transform(const float pos[3], const float q[4], const float scale[3], float out[16])
{
    float T[16], R[16], S[16], TR[16];
    /* glm::translate(I, p): column 3 = ((I0*px + I1*py) + I2*pz) + I3 */
    mat_identity(T);
    for (int row = 0; row < 4; ++row) {
        float i0 = row == 0 ? 1.0f : 0.0f, i1 = row == 1 ? 1.0f : 0.0f;
        float i2 = row == 2 ? 1.0f : 0.0f, i3 = row == 3 ? 1.0f : 0.0f;
        T[3 * 4 + row] = ((i0 * pos[0] + i1 * pos[1]) + i2 * pos[2]) + i3;
    }
    /* glm::mat4_cast(quat(w, x, y, z)) */
    float qx = q[0], qy = q[1], qz = q[2], qw = q[3];
    float qxx = qx * qx, qyy = qy * qy, qzz = qz * qz;
    float qxz = qx * qz, qxy = qx * qy, qyz = qy * qz;
    float qwx = qw * qx, qwy = qw * qy, qwz = qw * qz;
    mat_identity(R);
    R[0 * 4 + 0] = 1.0f - 2.0f * (qyy + qzz);
    R[0 * 4 + 1] = 2.0f * (qxy + qwz);
    R[0 * 4 + 2] = 2.0f * (qxz - qwy);
    R[1 * 4 + 0] = 2.0f * (qxy - qwz);
    R[1 * 4 + 1] = 1.0f - 2.0f * (qxx + qzz);
    R[1 * 4 + 2] = 2.0f * (qyz + qwx);
    R[2 * 4 + 0] = 2.0f * (qxz + qwy);
    R[2 * 4 + 1] = 2.0f * (qyz - qwx);
    R[2 * 4 + 2] = 1.0f - 2.0f * (qxx + qyy);
    /* glm::scale(I, s) */
    mat_identity(S);
    S[0] = scale[0]; S[5] = scale[1]; S[10] = scale[2];
    ora_mat4_mul(T, R, TR);
    ora_mat4_mul(TR, S, out);
}

void ora_flatten_sphere(const float G[16], const ora_sphere* in, ora_sphere* out)
{
    float c4[4] = {in->center[0], in->center[1], in->center[2], 1.0f}, tc[4];
    mat_vec4(G, c4, tc);
    *out = *in;
    out->center[0] = tc[0] / tc[3];
    out->center[1] = tc[1] / tc[3];
    out->center[2] = tc[2] / tc[3];
    float sx = v3len(v3make(G[0], G[1], G[2]));
    float sy = v3len(v3make(G[4], G[5], G[6]));
    float sz = v3len(v3make(G[8], G[9], G[10]));
    float uniform = (sx + sy + sz) / 3.0f;
    out->radius = in->radius * uniform;
}
