/*
 * atx.h -- C ABI of the MI355X-native Ataraxia path tracer (libatx.so).
 *
 * Plain C: opaque handles, plain pointers and sizes, integer status codes.
 * Every entry point returns ATX_OK (0) or a negative ATX_ERR_* code and never
 * exits the process; atx_last_error() returns a thread-local message for the
 * last failing call on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   atx_renderer_create / destroy        Renderer::Renderer / ~Renderer      Engine/include/Renderer.h:20-21
 *   atx_renderer_on_resize               Renderer::onResize                  Engine/src/Renderer.cu:98-146
 *   atx_renderer_set_scene               Renderer::allocateDeviceMemory      Engine/src/Renderer.cu:25-57
 *   atx_renderer_render                  Renderer::Render                    Engine/src/Renderer.cu:173-249
 *   atx_renderer_get_image               Renderer::getImage + Image pixels   Engine/include/Renderer.h:26, Renderer.cu:240-242
 *   atx_renderer_get/set_settings        Renderer::getSettings/setSettings   Engine/include/Renderer.h:28-29
 *   atx_renderer_reset_frame_index       Renderer::resetFrameIndex           Engine/include/Renderer.h:31
 *   atx_renderer_set_accel               Renderer::traceRay sphere loop      Engine/src/Renderer.cu:251-285
 *   atx_camera_ray_directions            Camera::UpdateRayDirection          Engine/src/Camera.cpp:161-195
 *   atx_scene_* (scene.json)             Utils::importScene / exportScene    Engine/src/Utils.cpp:53-62,175-187
 *
 * Struct layouts equal the reference's packed GLM structs so that arrays can
 * be handed over without conversion (Scene.h:11-56, SceneNode.h:11-21).
 */
#ifndef ATX_H
#define ATX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATX_OK 0
#define ATX_ERR_INVALID (-1)   /* bad argument or size mismatch            */
#define ATX_ERR_HIP (-2)       /* HIP runtime error (message has details) */
#define ATX_ERR_NOMEM (-3)     /* device or host allocation failed        */
#define ATX_ERR_STATE (-4)     /* call not valid in the current state     */
#define ATX_ERR_IO (-5)        /* file could not be read/written/parsed   */
#define ATX_ERR_BUDGET (-6)    /* request exceeds the HBM budget          */
#define ATX_ERR_NOTFOUND (-7)  /* unknown id                               */
#define ATX_ERR_NODEVICE (-8)  /* no HIP device available                  */

/* Sphere {vec3 center; float radius; int id}           SceneNode.h:11-21   (20 B) */
typedef struct atx_sphere { float center[3]; float radius; int32_t material; } atx_sphere;
/* Material {albedo, roughness, metallic, F0, emissionColor, emissionIntensity, id}
 *                                                        Scene.h:28-47      (52 B) */
typedef struct atx_material {
    float albedo[3];
    float roughness;
    float metallic;
    float F0[3];
    float emission_color[3];
    float emission_intensity;
    int32_t id;
} atx_material;
/* Light {vec3 position; vec3 color; float intensity}    Scene.h:17-26      (28 B) */
typedef struct atx_light { float position[3]; float color[3]; float intensity; } atx_light;
/* Settings {bool accumulation; bool skyLight; int maxBounces}  Scene.h:49-56  (8 B) */
typedef struct atx_settings {
    uint8_t accumulation;
    uint8_t sky_light;
    uint8_t reserved[2];
    int32_t max_bounces;
} atx_settings;
/* Camera state that determines the rays: position, direction, fov (degrees),
 * near/far clip and the viewport size (Camera.h:60-88).                      */
typedef struct atx_camera {
    float position[3];
    float direction[3];
    float fov_deg;
    float near_clip;
    float far_clip;
    uint32_t width;
    uint32_t height;
} atx_camera;

/* The layouts above are the ABI: checked at compile time in C and C++. */
#ifdef __cplusplus
#define ATX_STATIC_ASSERT(cond, msg) static_assert(cond, msg)
#else
#define ATX_STATIC_ASSERT(cond, msg) _Static_assert(cond, msg)
#endif
ATX_STATIC_ASSERT(sizeof(atx_sphere) == 20 && offsetof(atx_sphere, radius) == 12 &&
                  offsetof(atx_sphere, material) == 16, "atx_sphere must be the packed 20-byte Sphere");
ATX_STATIC_ASSERT(sizeof(atx_material) == 52 && offsetof(atx_material, roughness) == 12 &&
                  offsetof(atx_material, metallic) == 16 && offsetof(atx_material, F0) == 20 &&
                  offsetof(atx_material, emission_color) == 32 && offsetof(atx_material, emission_intensity) == 44 &&
                  offsetof(atx_material, id) == 48, "atx_material must be the packed 52-byte Material");
ATX_STATIC_ASSERT(sizeof(atx_light) == 28 && offsetof(atx_light, color) == 12 && offsetof(atx_light, intensity) == 24,
                  "atx_light must be the packed 28-byte Light");
ATX_STATIC_ASSERT(sizeof(atx_settings) == 8 && offsetof(atx_settings, sky_light) == 1 &&
                  offsetof(atx_settings, max_bounces) == 4, "atx_settings must be 8 bytes");
ATX_STATIC_ASSERT(sizeof(atx_camera) == 44 && offsetof(atx_camera, fov_deg) == 24 &&
                  offsetof(atx_camera, width) == 36, "atx_camera layout");

/* Accumulation buffer arithmetic.  NAIVE is the reference's `accum += c`
 * (Renderer.cu:165); KAHAN keeps a compensation term per channel.           */
#define ATX_ACCUM_NAIVE 0
#define ATX_ACCUM_KAHAN 1

/* Which spheres a ray tests (Renderer.cu:251-285 tests all of them).  BVH
 * (the default) walks a bounding-volume hierarchy built at scene upload and
 * gives the same closest hit, bit for bit, including the lowest index among
 * equal distances; NONE is the flat loop over every sphere.  Scenes with
 * fewer than 16 spheres always use the flat loop.                           */
#define ATX_ACCEL_BVH 0
#define ATX_ACCEL_NONE 1

/* Counters of the last render call. */
typedef struct atx_render_stats {
    uint64_t samples;        /* path samples (pixels x frames)                 */
    uint64_t rays;           /* rays of the rendered paths: primary/continuation
                                + shadow (as the reference traces them)         */
    uint64_t sphere_tests;   /* ray-sphere tests performed                      */
    uint64_t exact_retraces; /* traces resolved by the exact fallback loop      */
    uint64_t box_tests;      /* ray-box tests of BVH nodes                      */
    double gpu_ms;           /* HIP-event time of the render kernels            */
    uint64_t launches;       /* path-kernel launches                            */
    double path_ms;          /* HIP-event time of the path kernel alone, summed
                                over its launches (renderer; 0 for batches)     */
    uint64_t reused_rays;    /* of `rays`, those answered from the per-pixel
                                first-bounce cache (a primary ray and its shadow
                                ray are the same rays in every frame); the rest
                                were traced                                     */
} atx_render_stats;

const char* atx_last_error(void);
const char* atx_version(void);
int atx_device_count(int* count);
int atx_device_synchronize(void);

/* ---- Renderer: single scene, one frame per Render (the plain kernel) ----- */
typedef struct atx_renderer atx_renderer;

int atx_renderer_create(atx_renderer** out);
int atx_renderer_destroy(atx_renderer* r);
/* Allocates the RGBA8 image and accumulation for width x height and resets
 * the frame index to 1 when the size changes (Renderer.cu:98-146).          */
int atx_renderer_on_resize(atx_renderer* r, uint32_t width, uint32_t height);
/* Uploads a flattened scene (copied).  Material ids outside [0, m) are
 * clamped to 0, as Renderer.cu:30-37 does.  Does not reset the frame index. */
int atx_renderer_set_scene(atx_renderer* r, const atx_sphere* spheres, uint64_t num_spheres,
                           const atx_material* materials, uint64_t num_materials,
                           const atx_light* lights, uint64_t num_lights);
int atx_renderer_set_settings(atx_renderer* r, const atx_settings* s);
int atx_renderer_get_settings(const atx_renderer* r, atx_settings* s);
int atx_renderer_reset_frame_index(atx_renderer* r);
int atx_renderer_get_frame_index(const atx_renderer* r, uint32_t* frame_index);
int atx_renderer_set_accumulator(atx_renderer* r, int kind);
/* One frame: clears the accumulation when the frame index is 1, traces one
 * path per pixel, accumulates, writes RGBA8, then advances the frame index
 * (or keeps it at 1 when accumulation is off).  Returns after the GPU is done. */
int atx_renderer_render(atx_renderer* r, const atx_camera* cam);
/* `frames` consecutive Render calls queued back to back on the renderer's
 * stream with a single synchronisation at the end (no readback).           */
int atx_renderer_render_frames(atx_renderer* r, const atx_camera* cam, uint32_t frames);
/* Copies the RGBA8 image ((a<<24)|(b<<16)|(g<<8)|r per pixel) to the host. */
int atx_renderer_get_image(atx_renderer* r, uint32_t* rgba, uint64_t num_pixels);
/* Copies the accumulation buffer (4 floats per pixel: r, g, b, frames). */
int atx_renderer_get_accumulation(atx_renderer* r, float* rgba, uint64_t num_floats);
int atx_renderer_get_stats(const atx_renderer* r, atx_render_stats* stats);
/* One frame with per-pixel debug capture (parity tests): radiance of this
 * frame's sample (4 floats per pixel), closest-hit index of the continuation
 * ray and of the shadow ray of every bounce (-1 miss, -2 not traced), each
 * array width*height*max_bounces, pixel-major.  Any pointer may be NULL.     */
int atx_renderer_render_debug(atx_renderer* r, const atx_camera* cam, float* radiance,
                              int32_t* hit_ids, int32_t* shadow_ids, uint32_t max_bounces);

/* Kernel used by atx_renderer_render*: PLAIN = one lane per pixel, the
 * reference's per-pixel loop; WAVEFRONT (default) = the persistent
 * path-regenerating kernel (results are identical, only the scheduling
 * differs; it is faster also at one frame per call).                       */
#define ATX_KERNEL_PLAIN 0
#define ATX_KERNEL_WAVEFRONT 1
int atx_renderer_set_kernel(atx_renderer* r, int kind);
/* ATX_ACCEL_BVH (default) or ATX_ACCEL_NONE; results are identical.        */
int atx_renderer_set_accel(atx_renderer* r, int accel);
/* Allocates now the device buffers that render calls of up to `frames`
 * frames at the current size need (the persistent kernel's sample buffer,
 * primary directions, first-bounce cache once a scene is set), so that such
 * calls allocate nothing; a later call that needs more still grows them.
 * No counterpart in the reference (Renderer.cu:173-249 allocates and frees
 * per frame); for callers that time or pace render calls.                  */
int atx_renderer_reserve(atx_renderer* r, uint32_t frames);

/* ---- Allocation-free rendering ---------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops, harnesses that forbid hipMalloc): renders `frames`
 * frames of one scene with the plain kernel, exactly as
 * atx_renderer_render_frames does with ATX_KERNEL_PLAIN, frame indices
 * first_frame, first_frame + 1, ... (with accumulation off the index returns
 * to 1 after the first frame, Renderer.cu:245-248; frame 1 starts the
 * accumulation from zero).  The caller passes device buffers of width*height
 * entries: accum (4 floats per pixel), comp (4 floats per pixel, Kahan only,
 * else may be NULL) and image (RGBA8).  The scene (copied from host memory),
 * its sphere tree and the kernel's counters go into a fixed 8 MiB device
 * array that the library holds per device for the life of the process; a
 * scene that needs more returns ATX_ERR_INVALID.  No device or pinned host
 * memory is allocated.  Work is queued on `stream` (a hipStream_t, NULL =
 * the null stream) and the call returns when it has finished.  Calls on one
 * device share the array: calls that overlap (two threads, or two streams)
 * race.                                                                      */
int atx_render_frames_noalloc(const atx_sphere* spheres, uint64_t num_spheres,
                              const atx_material* materials, uint64_t num_materials,
                              const atx_light* lights, uint64_t num_lights, const atx_settings* settings,
                              const atx_camera* cam, uint32_t first_frame, uint32_t frames, int accel,
                              int accumulator, float* accum, float* comp, uint32_t* image, void* stream);

/* ---- Batch: many scenes, persistent launches of several frames each ------ */
/* No reference counterpart: the reference renders one scene per Render call.
 * Scene arrays are copied at add time; every scene keeps its own accumulation
 * and image on the device.                                                  */
typedef struct atx_scene_desc {
    const atx_sphere* spheres;
    uint64_t num_spheres;
    const atx_material* materials;
    uint64_t num_materials;
    const atx_light* lights;
    uint64_t num_lights;
    atx_settings settings;
    atx_camera camera;
    uint32_t spp;        /* frames (1 sample per pixel each) to accumulate */
    uint32_t reserved;
} atx_scene_desc;

typedef struct atx_batch_tuning {
    uint32_t chunk_px;          /* paths per work chunk: multiple of 64, <= 1024
                                   (0 = chosen per launch from its paths per wave)         */
    uint32_t workgroups;        /* persistent grid, 256-lane workgroups (0 = fill device) */
    uint32_t use_graph;         /* capture all launches in a HIP graph and replay it       */
    uint32_t frames_per_launch; /* frames each launch renders per scene (0 = 64)           */
    uint32_t waves_per_simd;    /* the path kernel's register budget: resident waves per
                                   SIMD it is compiled for, 5..8 (0 = 6 with trees, 7 for
                                   flat loops only); the launch bounds the autotuner picks */
} atx_batch_tuning;

typedef struct atx_batch atx_batch;

int atx_batch_create(atx_batch** out);
int atx_batch_destroy(atx_batch* b);
int atx_batch_add_scene(atx_batch* b, const atx_scene_desc* desc, uint32_t* scene_id);
int atx_batch_set_accumulator(atx_batch* b, int kind);
int atx_batch_set_accel(atx_batch* b, int accel);
int atx_batch_set_tuning(atx_batch* b, const atx_batch_tuning* t);
int atx_batch_get_tuning(const atx_batch* b, atx_batch_tuning* t);
/* Times candidate tunings on this batch (its first frames only) and applies
 * the fastest; with cache_path != NULL the choice is read from / written to
 * that file, keyed by device and workload signature.                       */
int atx_batch_autotune(atx_batch* b, const char* cache_path, atx_batch_tuning* chosen);
/* Renders every scene from frame 1 to its spp (accumulation restarts).      */
int atx_batch_render(atx_batch* b);
/* The same, limited to each scene's first max_frames frames.               */
int atx_batch_render_frames(atx_batch* b, uint32_t max_frames);
/* Per-pixel capture of the last rendered frame of one scene (parity tests):
 * enable before rendering, read after.  Layout as atx_renderer_render_debug. */
int atx_batch_capture(atx_batch* b, uint32_t scene_id, uint32_t max_bounces);
int atx_batch_get_capture(atx_batch* b, uint32_t scene_id, float* radiance, int32_t* hit_ids,
                          int32_t* shadow_ids);
int atx_batch_get_image(atx_batch* b, uint32_t scene_id, uint32_t* rgba, uint64_t num_pixels);
int atx_batch_get_accumulation(atx_batch* b, uint32_t scene_id, float* rgba, uint64_t num_floats);
int atx_batch_get_stats(const atx_batch* b, atx_render_stats* stats);

/* ---- Server: continuous-batching render service ----------------------------
 * No reference counterpart.  A worker thread renders all admitted requests
 * (a scene + spp) together, frames_per_tick frames per request per launch of
 * the persistent kernel, and admits queued requests between launches while
 * the admitted ones fill less than tick_mpaths million paths per launch and
 * their device buffers fit the HBM budget.  Among queued requests the one
 * with the highest response ratio (waiting time + estimated render time) /
 * estimated render time goes first: short requests overtake long ones, and a
 * long one's priority grows while it waits, so none starves.  Every admitted
 * request stays device-resident until it completes; its image and
 * accumulation are then copied to host memory and its buffers released.    */
#define ATX_REQ_QUEUED 0
#define ATX_REQ_RUNNING 1
#define ATX_REQ_DONE 2
#define ATX_REQ_FAILED 3

typedef struct atx_server_config {
    uint64_t hbm_budget_bytes;  /* device bytes for request buffers (0 = half of free HBM) */
    uint32_t frames_per_tick;   /* frames each request advances per launch (0 = 64)         */
    uint32_t chunk_px;          /* paths per work chunk (0 = chosen per launch)             */
    uint32_t max_active;        /* cap on resident requests (0 = budget only)               */
    uint32_t kahan;             /* 1 = compensated accumulation                             */
    uint32_t accel;             /* ATX_ACCEL_BVH (0) or ATX_ACCEL_NONE                      */
    uint32_t tick_mpaths;       /* admit until a launch holds this many million paths
                                   (0 = 128; UINT32_MAX = admit whatever fits the budget)   */
} atx_server_config;

typedef struct atx_server_stats {
    uint64_t submitted, admitted, completed, rejected, failed;
    uint64_t queued, active;
    uint64_t samples;              /* path samples of completed requests        */
    uint64_t ticks;                /* launches issued                           */
    uint64_t hbm_in_use;           /* device bytes admitted requests hold (block
                                      capacities + share of the tick's samples) */
    uint64_t hbm_high_water;       /* max of hbm_in_use                         */
    uint64_t device_used_high_water; /* max device memory in use (hipMemGetInfo) */
    double gpu_busy_ms;            /* HIP-event time of all launches            */
    double admit_ms;               /* worker time admitting requests (uploads)  */
    double retire_ms;              /* time retiring them (result copies), on a
                                      thread of its own                         */
    double wait_ms;                /* worker time waiting for launches to end   */
    double release_ms;             /* worker time releasing finished requests'
                                      device buffers (between launches)         */
    double gap_ms;                 /* device time between the end of a launch
                                      and the start of the next one, summed     */
} atx_server_stats;

typedef struct atx_request_info {
    int32_t state;                 /* ATX_REQ_*                                 */
    uint32_t frames_done;
    double t_submit_s, t_admit_s, t_done_s;  /* seconds since server creation   */
} atx_request_info;

typedef struct atx_server atx_server;

int atx_server_create(atx_server** out, const atx_server_config* cfg);
/* Stops the worker after the launch in flight and frees every buffer.       */
int atx_server_destroy(atx_server* s);
/* Copies the request; ATX_ERR_BUDGET if it could never fit the budget.      */
int atx_server_submit(atx_server* s, const atx_scene_desc* req, uint64_t* request_id);
int atx_server_get_info(atx_server* s, uint64_t request_id, atx_request_info* info);
/* Waits until the request is DONE or FAILED (timeout_s < 0: no limit).       */
int atx_server_wait(atx_server* s, uint64_t request_id, double timeout_s);
/* Copies a finished request's RGBA8 image and/or accumulation (either may be
 * NULL); release != 0 drops the host copy afterwards.                        */
int atx_server_get_result(atx_server* s, uint64_t request_id, uint32_t* rgba, uint64_t num_pixels,
                          float* accum, uint64_t num_floats, int release);
int atx_server_get_stats(atx_server* s, atx_server_stats* stats);

/* ---- Scene graph and scene.json (Scene.h:58-80, SceneNode.h:23-66) --------
 * A host-side scene: a tree of nodes with translate * rotate(quaternion) *
 * scale transforms holding spheres, plus materials, lights, settings and the
 * camera's position, direction and fov.  Node 0 is the root ("Scene").      */
typedef struct atx_scene_graph atx_scene_graph;

int atx_scene_create(atx_scene_graph** out);
int atx_scene_destroy(atx_scene_graph* g);
/* Utils::importScene (Utils.cpp:175-187): a file that cannot be opened gives
 * an empty default scene (ATX_OK); malformed JSON gives ATX_ERR_IO.          */
int atx_scene_load_json(const char* path, atx_scene_graph** out);
int atx_scene_parse_json(const char* text, uint64_t len, atx_scene_graph** out);
/* Utils::exportScene (Utils.cpp:53-62): 4-space indented JSON, sorted keys. */
int atx_scene_save_json(const atx_scene_graph* g, const char* path);
/* Writes the JSON text (NUL-terminated) when buf != NULL; *len = its length. */
int atx_scene_to_json(const atx_scene_graph* g, char* buf, uint64_t cap, uint64_t* len);
int atx_scene_add_node(atx_scene_graph* g, uint32_t parent, const char* name, const float* position,
                       const float* rotation_xyzw, const float* scale, uint32_t* node);
int atx_scene_set_node_transform(atx_scene_graph* g, uint32_t node, const float* position,
                                 const float* rotation_xyzw, const float* scale);
int atx_scene_add_sphere(atx_scene_graph* g, uint32_t node, const atx_sphere* s);
int atx_scene_add_material(atx_scene_graph* g, const atx_material* m);
int atx_scene_add_light(atx_scene_graph* g, const atx_light* l);
int atx_scene_set_settings(atx_scene_graph* g, const atx_settings* s);
int atx_scene_get_settings(const atx_scene_graph* g, atx_settings* s);
int atx_scene_set_camera(atx_scene_graph* g, const float* position, const float* direction, float fov_deg);
int atx_scene_get_camera(const atx_scene_graph* g, float* position, float* direction, float* fov_deg);
int atx_scene_counts(const atx_scene_graph* g, uint64_t* nodes, uint64_t* spheres, uint64_t* materials,
                     uint64_t* lights);
/* Node `node` (ids are assigned parents-first, siblings in order): name
 * (NUL-terminated, truncated to name_cap), transform, parent (0 for the root)
 * and number of spheres.  Any output pointer may be NULL.                   */
int atx_scene_get_node(const atx_scene_graph* g, uint32_t node, char* name, uint64_t name_cap, float* position,
                       float* rotation_xyzw, float* scale, uint32_t* parent, uint64_t* num_spheres);
int atx_scene_get_node_spheres(const atx_scene_graph* g, uint32_t node, atx_sphere* out, uint64_t cap,
                               uint64_t* count);
/* Renderer::traverseSceneGraph (Renderer.cu:67-96): world-space spheres in
 * pre-order (a node's spheres, then its children), centre = M*(c,1) / w,
 * radius *= mean of the lengths of M's first three columns.  With out ==
 * NULL only *count is set.                                                   */
int atx_scene_flatten(const atx_scene_graph* g, atx_sphere* out, uint64_t cap, uint64_t* count);
int atx_scene_get_materials(const atx_scene_graph* g, atx_material* out, uint64_t cap, uint64_t* count);
int atx_scene_get_lights(const atx_scene_graph* g, atx_light* out, uint64_t cap, uint64_t* count);

/* ---- Headless image output (replaces the Vulkan Image, Image.cpp:183-271) -- */
/* Writes RGBA8 pixels ((a<<24)|(b<<16)|(g<<8)|r) as a binary PPM (P6, alpha
 * dropped).  Row 0 of the buffer is the bottom of the view (clip y = -1);
 * flip_vertical != 0 writes it last so that the file reads top-down.        */
int atx_write_ppm(const char* path, const uint32_t* rgba, uint32_t width, uint32_t height, int flip_vertical);

/* ---- Camera: host-side ray directions (Camera::UpdateRayDirection) -------- */
/* Writes width*height*3 floats, bit-identical to what the kernels generate. */
int atx_camera_ray_directions(const atx_camera* cam, float* dirs, uint64_t num_floats);
/* Inverse projection and inverse view matrices (column-major, 16 floats each). */
int atx_camera_matrices(const atx_camera* cam, float* inv_proj, float* inv_view);

#ifdef __cplusplus
}
#endif
#endif
