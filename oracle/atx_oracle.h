/*
 * atx_oracle.h -- CPU restatement of Ataraxia's per-pixel path tracer.
 *
 * TEST INFRASTRUCTURE ONLY.  Results are compared against this library; it
 * may be loaded by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py, and by nothing else.  The product (ataraxia_amd/) never
 * links or calls it.
 *
 * Every function restates one reference function in plain scalar C99 with
 * IEEE-754 binary32 arithmetic, no FMA contraction (-ffp-contract=off), and
 * correctly rounded division and square root.  GLM semantics (the reference's
 * unvendored vector library) are restated explicitly:
 *   dot(a,b)      = (a.x*b.x + a.y*b.y) + a.z*b.z
 *   normalize(v)  = v * (1 / sqrt(dot(v,v)))
 *   max(a,b)      = a < b ? b : a        min(a,b) = b < a ? b : a
 *   mix(x,y,a)    = x*(1-a) + y*a
 *   mat4*vec4     = (m0*x + m1*y) + (m2*z + m3*w)
 *
 * sin, cos and pow(x,5) are this project's own implementations (Cody-Waite
 * reduction by pi/2 in three parts with fused multiply-add, Taylor polynomials
 * on [-pi/4, pi/4], and x^5 = (x^2)^2 * x); the device code evaluates the very
 * same sequence so that hit indices can match bit for bit.
 *
 * Parity status of each function is listed in the header of atx_oracle.c.
 */
#ifndef ATX_ORACLE_H
#define ATX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Layouts match the reference's packed GLM structs (Scene.h, SceneNode.h). */
typedef struct { float center[3]; float radius; int32_t material; } ora_sphere;            /* 20 B */
typedef struct {
    float albedo[3]; float roughness; float metallic; float F0[3];
    float emission_color[3]; float emission_intensity; int32_t id;
} ora_material;                                                                             /* 52 B */
typedef struct { float position[3]; float color[3]; float intensity; } ora_light;          /* 28 B */
typedef struct { uint8_t accumulation; uint8_t sky_light; uint8_t pad[2]; int32_t max_bounces; } ora_settings; /* 8 B */
typedef struct {
    float position[3]; float direction[3];
    float fov_deg; float near_clip; float far_clip;
    uint32_t width; uint32_t height;
} ora_camera;

typedef struct {
    const ora_sphere* spheres; uint64_t num_spheres;
    const ora_material* materials; uint64_t num_materials;
    const ora_light* lights; uint64_t num_lights;
} ora_scene;

/* ---- primitives ---------------------------------------------------------- */
uint32_t ora_pcg_hash(uint32_t seed);
float    ora_pcg_float(uint32_t* seed);
float    ora_sin(float x);
float    ora_cos(float x);
float    ora_pow5(float x);

/* ---- camera (host ray generation) ---------------------------------------- */
/* Column-major 4x4 matrices, m[col*4 + row], as GLM stores them. */
void ora_perspective(float fov_deg, float aspect, float znear, float zfar, float out[16]);
void ora_look_at(const float eye[3], const float center[3], const float up[3], float out[16]);
void ora_inverse(const float m[16], float out[16]);
void ora_camera_matrices(const ora_camera* cam, float inv_proj[16], float inv_view[16]);
void ora_ray_direction(const float inv_proj[16], const float inv_view[16],
                       uint32_t x, uint32_t y, uint32_t width, uint32_t height, float out[3]);
void ora_ray_directions(const ora_camera* cam, float* dirs /* width*height*3 */);

/* ---- intersection and shading -------------------------------------------- */
/* Returns the closest sphere index or -1; *t_out gets tmin (FLT_MAX on miss). */
int32_t ora_trace_ray(const float origin[3], const float dir[3],
                      const ora_sphere* spheres, uint64_t n, float* t_out);
void ora_cook_torrance(const float albedo[3], const float F0[3], float metallic, float roughness,
                       const float N[3], const float V[3], const float L[3], float out[3]);
void ora_sample_cosine(const float N[3], uint32_t* seed, float out[3]);
void ora_sample_ggx(const float N[3], float roughness, uint32_t* seed, float out[3]);

/* Per-path debug record.  hit_ids[b] is the closest-hit index of the bounce-b
 * continuation ray (primary for b == 0) and shadow_ids[b] that of its shadow ray
 * (-1 = miss, -2 = not traced).  Arrays hold max_bounces entries (rr: 2 per
 * bounce, dirs: 3 per bounce).  Every array pointer may be NULL.  The last
 * four expose the random stream and the continuation decisions per bounce so
 * that an independent restatement can check them:
 *   seeds[b]     the u32 seed right after `seed += b` (Renderer.cu:306)
 *   light_ids[b] PcgHash(seed) % numLights, -1 without lights (:340)
 *   rr[2b..2b+1] Russian roulette p and u (:375-376); NaN if not reached
 *   dirs[3b..]   direction of the ray traced at bounce b (primary at b = 0);
 *                NaN if not reached                                       */
typedef struct {
    int32_t* hit_ids;
    int32_t* shadow_ids;
    uint64_t rays;        /* rays traced by this path (continuation + shadow) */
    int32_t  bounces;     /* bounce iterations entered                       */
    uint32_t* seeds;
    int32_t*  light_ids;
    float*    rr;
    float*    dirs;
} ora_path_debug;

void ora_per_pixel(uint32_t x, uint32_t y, uint32_t width, const float dir[3],
                   const float cam_pos[3], const ora_scene* scene, uint32_t frame_index,
                   const ora_settings* settings, float out_rgba[4], ora_path_debug* dbg);

/* ---- frames --------------------------------------------------------------- */
/* One Renderer::Render frame over rows [row0,row1): accum (float4 per pixel) +=
 * radiance, image = RGBA8(clamp(accum / frame_index)).  When comp != NULL the
 * accumulation is Kahan-compensated (sum in accum, compensation in comp).
 * When radiance != NULL the per-sample rgba is also written there.
 * rays (optional) receives the number of rays traced.                        */
void ora_render_frame(const ora_camera* cam, const ora_scene* scene, const ora_settings* settings,
                      uint32_t frame_index, float* accum, float* comp, uint32_t* image,
                      float* radiance, uint32_t row0, uint32_t row1, uint64_t* rays);

/* The same over all rows with `threads` POSIX threads, `frames` consecutive
 * frames starting at first_frame (frame_index 1 clears accum/comp first).    */
int  ora_render_frames_mt(const ora_camera* cam, const ora_scene* scene, const ora_settings* settings,
                          uint32_t first_frame, uint32_t frames, float* accum, float* comp,
                          uint32_t* image, int threads, uint64_t* rays);

uint32_t ora_rgba8(const float rgba[4]);

/* Accumulation of `frames` consecutive frames (first_frame, first_frame+1,
 * ...; the reference's frame order) for a list of pixels (x + y*width),
 * as Renderer::Render leaves it: accum[4*i..] for pixels[i], the naive sum
 * of each frame's sample, or the Kahan-compensated one when kahan != 0.
 * image (optional) gets RGBA8 of accum / last frame index.  Runs on
 * `threads` POSIX threads over the pixel list.
 * ora_accumulate_pixels_f64 sums the same float32 samples in float64 (the
 * precision reference; accum holds 4 doubles per pixel).                  */
int  ora_accumulate_pixels(const ora_camera* cam, const ora_scene* scene, const ora_settings* settings,
                           uint32_t first_frame, uint32_t frames, const uint32_t* pixels, uint64_t n,
                           int kahan, float* accum, uint32_t* image, int threads);
int  ora_accumulate_pixels_f64(const ora_camera* cam, const ora_scene* scene, const ora_settings* settings,
                               uint32_t first_frame, uint32_t frames, const uint32_t* pixels, uint64_t n,
                               double* accum, int threads);

/* Per-pixel debug capture of one frame over rows [row0,row1): radiance (4
 * floats per pixel), continuation and shadow hit ids per bounce (pixel-major,
 * max_bounces per pixel, -2 = not traced), rays per pixel.  Any output may be
 * NULL.                                                                       */
void ora_render_frame_debug(const ora_camera* cam, const ora_scene* scene, const ora_settings* settings,
                            uint32_t frame_index, float* radiance, int32_t* hit_ids, int32_t* shadow_ids,
                            uint32_t max_bounces, uint32_t* rays, uint32_t row0, uint32_t row1);

/* ---- scene-graph flattening ---------------------------------------------- */
/* Node transform = translate(p) * mat4_cast(q) * scale(s); q = (x,y,z,w).     */
void ora_node_local_transform(const float pos[3], const float quat_xyzw[4], const float scale[3], float out[16]);
void ora_mat4_mul(const float a[16], const float b[16], float out[16]);
void ora_flatten_sphere(const float global[16], const ora_sphere* in, ora_sphere* out);

#ifdef __cplusplus
}
#endif
#endif
