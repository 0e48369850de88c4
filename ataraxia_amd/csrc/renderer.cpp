// renderer.cpp -- the single-scene Renderer behind the C ABI.
//
// Mirrors Renderer (Engine/include/Renderer.h:18-66, Engine/src/Renderer.cu):
// the handle owns the device scene, the accumulation buffer(s), the RGBA8
// image, the settings and the frame index.  Differences by design:
//  * rays are generated on the device from the camera's inverse matrices,
//    so there is no per-frame W*H*12 B upload, malloc or free
//    (Camera.cpp:197-220);
//  * no per-frame occupancy/device-property queries or full-device syncs;
//    work is queued on the handle's own stream and `render` waits on it;
//  * the RGBA8 image stays on the device until atx_renderer_get_image;
//  * the persistent kernel (the default) renders up to 256 frames per
//    launch, and keeps per camera and scene the primary directions and the
//    first-bounce cache (built for calls of two or more frames, or once the
//    camera has not moved since the previous call).
#include <algorithm>
#include <new>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "camera_host.h"
#include "device_types.h"
#include "host_common.h"
#include "wavefront_host.h"

namespace atx {

struct Renderer {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<hipEvent_t> wf_ev;   // per wavefront launch: path kernel start, end
    DevScene geo;
    bool has_scene = false;
    int kernel = ATX_KERNEL_WAVEFRONT;
    int accel = ATX_ACCEL_BVH;
    DevBuf wf_table, wf_queue, wf_samples, wf_dirs, wf_first;
    // wf_dirs holds the primary rays of dirs_cam; wf_first the first-bounce
    // cache of that camera and the current scene (set_scene clears first_valid)
    atx_camera dirs_cam{}, prev_cam{};
    bool dirs_valid = false, first_valid = false, prev_valid = false;
    // pixels of wf_first's hit list, read back after the call that built it
    // (n_hit_exact: for the current cache; else the previous cache's)
    uint32_t n_hit = 0;
    bool n_hit_exact = false;
    DevBuf dbg_rad, dbg_hit, dbg_sh;   // debug capture buffers, reused across calls
    uint32_t wf_chunk = 0;   // jobs per chunk; 0: auto_chunk_px per launch
    uint32_t wf_frames = kRendererFramesPerLaunch;
    uint32_t wf_waves = 0;   // path-kernel instance (resident waves per SIMD); 0: default
    DevBuf accum, comp, image, counters;
    // pinned host memory the call's statistics are copied into (counters,
    // then the hit list's length): copies into pageable memory block the
    // host until the stream is idle, one after the other
    unsigned long long* readback = nullptr;
    uint32_t width = 0, height = 0;
    atx_settings settings{1, 0, {0, 0}, 15};  // Scene.h:49-56 defaults
    uint32_t frame_index = 1;
    int accumulator = ATX_ACCUM_NAIVE;
    size_t lds_budget = 64 * 1024;
    bool exact_loop = false;
    atx_render_stats stats{};
};

static int renderer_launch_frames(Renderer* r, const atx_camera* cam, uint32_t frames, float* dbg_rad,
                                  int32_t* dbg_hit, int32_t* dbg_shadow, uint32_t dbg_bounces)
{
    ATX_CHECK_ARG(cam, "camera is NULL");
    if (r->width == 0 || r->height == 0) {
        set_error("renderer has no viewport: call atx_renderer_on_resize first");
        return ATX_ERR_STATE;
    }
    ATX_CHECK_ARG(cam->width == r->width && cam->height == r->height,
                  "camera viewport %ux%u does not match renderer %ux%u", cam->width, cam->height, r->width,
                  r->height);
    if (!r->has_scene) {
        set_error("no scene: call atx_renderer_set_scene first");
        return ATX_ERR_STATE;
    }
    ATX_CHECK_ARG(r->settings.max_bounces >= 0, "max_bounces must be >= 0");

    PlainArgs a{};
    a.sc.sph = r->geo.sph.as<float4>();
    a.sc.sph_mat = r->geo.sph_mat.as<int32_t>();
    a.sc.mat = r->geo.mat.as<float4>();
    a.sc.light = r->geo.light.as<float4>();
    a.sc.n_sph = r->geo.n_sph;
    a.sc.n_mat = r->geo.n_mat;
    a.sc.n_light = r->geo.n_light;
    const int accel = r->exact_loop ? ATX_ACCEL_NONE : r->accel;
    a.sc.bvh = r->geo.accel(accel);
    camera_inverse_matrices(cam->position, cam->direction, cam->fov_deg, cam->near_clip, cam->far_clip,
                            cam->width, cam->height, a.cam.inv_proj, a.cam.inv_view);
    memcpy(a.cam.pos, cam->position, sizeof(a.cam.pos));
    a.cam.width = r->width;
    a.cam.height = r->height;
    a.max_bounces = r->settings.max_bounces;
    a.sky_light = r->settings.sky_light ? 1u : 0u;
    a.accum = r->accum.as<float4>();
    a.comp = r->accumulator == ATX_ACCUM_KAHAN ? r->comp.as<float4>() : nullptr;
    a.image = r->image.as<uint32_t>();
    a.counters = r->counters.as<unsigned long long>();

    const size_t npx = (size_t)r->width * r->height;
    DevBuf& d_rad = r->dbg_rad;
    DevBuf& d_hit = r->dbg_hit;
    DevBuf& d_sh = r->dbg_sh;
    uint32_t variant = 0;
    if ((size_t)r->geo.n_sph * sizeof(float4) <= r->lds_budget) variant |= kVarLds;
    if (r->accumulator == ATX_ACCUM_KAHAN) variant |= kVarKahan;
    if (r->exact_loop) variant |= kVarExactLoop;
    if (a.sc.bvh.n_nodes) variant = (variant & ~(uint32_t)kVarLds) | kVarBvh;
    if (dbg_rad || dbg_hit || dbg_shadow) {
        variant |= kVarDebug;
        if (dbg_rad) { ATX_HIP(d_rad.reserve(npx * sizeof(float4))); a.dbg_radiance = d_rad.as<float4>(); }
        if (dbg_hit) { ATX_HIP(d_hit.reserve(npx * dbg_bounces * 4)); a.dbg_hit = d_hit.as<int32_t>(); }
        if (dbg_shadow) { ATX_HIP(d_sh.reserve(npx * dbg_bounces * 4)); a.dbg_shadow = d_sh.as<int32_t>(); }
        a.dbg_bounces = dbg_bounces;
    }

    // wavefront mode: launches of up to wf_frames frames each (one frame per
    // launch when accumulation is off: the frame index then stays 1)
    std::vector<WfScene> table;
    std::vector<uint32_t> launch_frames, launch_chunk;
    const bool wf = r->kernel == ATX_KERNEL_WAVEFRONT && frames > 0;
    bool use_first = false;
    WfLaunch shape{0, false};
    if (wf) {
        uint32_t fi = r->frame_index, left = frames;
        const uint32_t fit = (uint32_t)std::max<size_t>(1, kSampleBufferBytes / (npx * kSampleBytes));
        const uint32_t per_launch = std::min(r->wf_frames, fit);
        while (left > 0) {
            const uint32_t k = r->settings.accumulation ? (left < per_launch ? left : per_launch) : 1u;
            WfScene e;
            fill_wf_scene(r->geo, r->settings, *cam, a.accum, a.comp, a.image, a.comp != nullptr, fi, k, accel,
                          &e);
            e.dbg_radiance = a.dbg_radiance;
            e.dbg_hit = a.dbg_hit;
            e.dbg_shadow = a.dbg_shadow;
            e.dbg_bounces = a.dbg_bounces;
            table.push_back(e);
            launch_frames.push_back(k);
            fi = r->settings.accumulation ? fi + k : 1u;
            left -= k;
        }
        // every launch runs after the previous one on the stream: one sample buffer serves all
        size_t most = 0;
        for (const WfScene& e : table) most = std::max(most, launch_samples(&e, 1));
        ATX_HIP(r->wf_samples.reserve(most * kSampleBytes));
        shape.flat_only = wf_flat_only(table.data(), table.size());
        // the instance: the caller's, else by the jobs of the largest launch
        // (every pixel in every frame, an upper bound of the hit list's)
        shape.waves = r->wf_waves ? r->wf_waves
                      : shape.flat_only ? 0u
                                        : wf_tree_waves((uint64_t)npx * launch_frames[0]);
        shape.workgroups = wavefront_full_grid(shape.flat_only, shape.waves);
        // The first-bounce cache pays from two frames per camera on: a call
        // of one frame uses it once the camera has stayed the same since the
        // previous call (a camera that moves every frame never builds it).
        use_first = frames >= 2 || (r->prev_valid && memcmp(&r->prev_cam, cam, sizeof(*cam)) == 0);
        r->prev_cam = *cam;
        r->prev_valid = true;
        ATX_HIP(r->wf_dirs.reserve(npx * sizeof(float4)));
        if (use_first) ATX_HIP(r->wf_first.reserve(first_cache_bytes(npx, r->geo.n_light)));
        for (WfScene& e : table) {
            e.dirs = r->wf_dirs.as<float4>();
            if (use_first) set_first_cache(e, r->wf_first.get());
        }
        // primary rays once per camera, the first-bounce cache (and its hit
        // count, which the layout needs) once per camera and scene
        if (!r->dirs_valid || memcmp(&r->dirs_cam, cam, sizeof(*cam)) != 0) {
            r->dirs_valid = r->first_valid = false;
            ATX_HIP(launch_camera_dirs(table[0], r->wf_dirs.as<float4>(), r->stream));
            r->dirs_cam = *cam;
            r->dirs_valid = true;
        }
        // The path kernel reads the hit list's length on the device, so the
        // launches follow k_first_hit without the host waiting for it: until
        // the length is read back (with this call's statistics) the chunks
        // are laid out for every pixel and their size chosen for the previous
        // cache's length.
        if (use_first && !r->first_valid) {
            ATX_HIP(launch_first_hits(table[0], r->stream));
            r->first_valid = true;
            r->n_hit_exact = false;
        }
        for (WfScene& e : table) {
            if (use_first) use_hit_list(e, r->wf_first.get(), r->n_hit_exact ? r->n_hit : (uint32_t)npx);
            const uint64_t est =
                e.hit_px && !r->n_hit_exact && r->n_hit ? (uint64_t)r->n_hit * e.frames : scene_jobs(e);
            const uint32_t P = r->wf_chunk ? r->wf_chunk : auto_chunk_px(est, shape.workgroups);
            launch_chunk.push_back(P);
            WfPass lp{};
            if (!layout_launch(&e, 1, P, r->wf_samples.as<float3>(), &lp)) {
                set_error("render: %u x %u pixels x %u frames exceed one launch", e.width, e.height, e.frames);
                return ATX_ERR_INVALID;
            }
        }
        ATX_HIP(r->wf_table.reserve(table.size() * sizeof(WfScene)));
        ATX_HIP(r->wf_queue.reserve(table.size() * sizeof(unsigned int)));
        ATX_HIP(hipMemcpyAsync(r->wf_table.get(), table.data(), table.size() * sizeof(WfScene),
                               hipMemcpyHostToDevice, r->stream));
        while (r->wf_ev.size() < 2 * table.size()) {
            hipEvent_t ev = nullptr;
            ATX_HIP(hipEventCreate(&ev));
            r->wf_ev.push_back(ev);
        }
    }

    ATX_HIP(hipMemsetAsync(r->counters.get(), 0, kCtrCount * sizeof(unsigned long long), r->stream));
    if (wf) ATX_HIP(launch_zero_u32(r->wf_queue.as<unsigned int>(), (uint32_t)table.size(), r->stream));
    ATX_HIP(hipEventRecord(r->ev0, r->stream));
    if (wf) {
        for (size_t l = 0; l < table.size(); ++l) {
            WfPass wp{};
            wp.scenes = r->wf_table.as<WfScene>() + l;
            wp.n_scenes = 1;
            wp.total_chunks = table[l].n_chunks;
            wp.chunk_px = launch_chunk[l];
            wp.total_px = table[l].width * table[l].height;
            wp.queue = r->wf_queue.as<unsigned int>() + l;
            wp.counters = a.counters;
            ATX_HIP(hipEventRecord(r->wf_ev[2 * l], r->stream));
            ATX_HIP(launch_wavefront(wp, shape, r->stream, r->wf_ev[2 * l + 1]));
            // Renderer.cu:245-248, once per frame of the launch
            for (uint32_t f = 0; f < launch_frames[l]; ++f) {
                if (r->settings.accumulation) r->frame_index++;
                else r->frame_index = 1;
            }
        }
    } else {
        for (uint32_t f = 0; f < frames; ++f) {
            a.frame = r->frame_index;
            ATX_HIP(launch_plain(a, variant, r->stream));
            // Renderer.cu:245-248
            if (r->settings.accumulation) r->frame_index++;
            else r->frame_index = 1;
        }
    }
    ATX_HIP(hipEventRecord(r->ev1, r->stream));
    unsigned long long* ctr = r->readback;
    ATX_HIP(hipMemcpyAsync(ctr, r->counters.get(), kCtrCount * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           r->stream));
    const bool read_hits = use_first && !r->n_hit_exact;
    uint32_t* n_hit_read = reinterpret_cast<uint32_t*>(r->readback + kCtrCount);
    if (read_hits)
        ATX_HIP(hipMemcpyAsync(n_hit_read, first_cache_view(r->wf_first.get(), npx, r->geo.n_light).n_hits,
                               sizeof(uint32_t), hipMemcpyDeviceToHost, r->stream));
    if (dbg_rad) ATX_HIP(hipMemcpyAsync(dbg_rad, d_rad.get(), npx * sizeof(float4), hipMemcpyDeviceToHost, r->stream));
    if (dbg_hit) ATX_HIP(hipMemcpyAsync(dbg_hit, d_hit.get(), npx * dbg_bounces * 4, hipMemcpyDeviceToHost, r->stream));
    if (dbg_shadow) ATX_HIP(hipMemcpyAsync(dbg_shadow, d_sh.get(), npx * dbg_bounces * 4, hipMemcpyDeviceToHost, r->stream));
    ATX_HIP(hipStreamSynchronize(r->stream));
    float ms = 0.0f;
    ATX_HIP(hipEventElapsedTime(&ms, r->ev0, r->ev1));
    double path_ms = ms;
    if (wf) {
        path_ms = 0.0;
        for (size_t l = 0; l < table.size(); ++l) {
            float t = 0.0f;
            ATX_HIP(hipEventElapsedTime(&t, r->wf_ev[2 * l], r->wf_ev[2 * l + 1]));
            path_ms += t;
        }
    }
    if (read_hits) {
        r->n_hit = *n_hit_read;
        r->n_hit_exact = true;
    }
    for (WfScene& e : table)
        if (e.hit_px) e.n_hit = r->n_hit;
    r->stats.samples = (uint64_t)npx * frames;
    const uint64_t skipped = wf ? skipped_primaries(table.data(), table.size()) : 0;
    r->stats.rays = ctr[kCtrRays] + ctr[kCtrReused] + skipped;
    r->stats.reused_rays = ctr[kCtrReused] + skipped;
    r->stats.sphere_tests = ctr[kCtrTests];
    r->stats.exact_retraces = ctr[kCtrExact];
    r->stats.box_tests = ctr[kCtrBoxes];
    r->stats.gpu_ms = ms;
    r->stats.launches = wf ? table.size() : frames;
    r->stats.path_ms = path_ms;
    return ATX_OK;
}

}  // namespace atx

using namespace atx;

extern "C" {

int atx_renderer_create(atx_renderer** out)
{
    ATX_CHECK_ARG(out, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        set_error("no HIP device available");
        return ATX_ERR_NODEVICE;
    }
    Renderer* r = new (std::nothrow) Renderer();
    if (!r) { set_error("out of host memory"); return ATX_ERR_NOMEM; }
    hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&r->ev0);
    if (e == hipSuccess) e = hipEventCreate(&r->ev1);
    if (e == hipSuccess) e = r->counters.reserve(kCtrCount * sizeof(unsigned long long));
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&r->readback, (kCtrCount + 1) * sizeof(unsigned long long), hipHostMallocDefault);
    if (e != hipSuccess) {
        set_error("renderer create: %s", hipGetErrorString(e));
        atx_renderer_destroy(reinterpret_cast<atx_renderer*>(r));
        return ATX_ERR_HIP;
    }
    if (const char* env = getenv("ATX_EXACT_LOOP")) r->exact_loop = atoi(env) != 0;
    if (const char* env = getenv("ATX_LDS_BUDGET")) r->lds_budget = (size_t)strtoull(env, nullptr, 10);
    if (const char* env = getenv("ATX_WF_CHUNK")) {
        const uint32_t c = (uint32_t)strtoul(env, nullptr, 10);
        if (c >= 64 && c % 64 == 0 && c <= kMaxChunkPx) r->wf_chunk = c;
    }
    if (const char* env = getenv("ATX_WF_WAVES")) {
        const uint32_t w = (uint32_t)strtoul(env, nullptr, 10);
        if (w >= kWfWavesMin && w <= kWfWavesMax) r->wf_waves = w;
    }
    if (const char* env = getenv("ATX_WF_FRAMES")) {
        const uint32_t f = (uint32_t)strtoul(env, nullptr, 10);
        if (f >= 1 && f <= 4096) r->wf_frames = f;
    }
    *out = reinterpret_cast<atx_renderer*>(r);
    return ATX_OK;
}

int atx_renderer_destroy(atx_renderer* h)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    if (!r) return ATX_OK;
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    if (r->ev0) (void)hipEventDestroy(r->ev0);
    if (r->ev1) (void)hipEventDestroy(r->ev1);
    for (hipEvent_t ev : r->wf_ev) (void)hipEventDestroy(ev);
    if (r->stream) (void)hipStreamDestroy(r->stream);
    if (r->readback) (void)hipHostFree(r->readback);
    delete r;
    return ATX_OK;
}

int atx_renderer_on_resize(atx_renderer* h, uint32_t width, uint32_t height)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r, "renderer is NULL");
    ATX_CHECK_ARG(width > 0 && height > 0, "Width or height cannot be zero");
    if (r->width == width && r->height == height) return ATX_OK;  // Renderer.cu:100-101
    const size_t npx = (size_t)width * height;
    ATX_HIP(hipStreamSynchronize(r->stream));
    r->image.release(); r->accum.release(); r->comp.release();
    r->dirs_valid = r->first_valid = r->prev_valid = false;
    ATX_HIP(r->image.reserve(npx * sizeof(uint32_t)));
    ATX_HIP(r->accum.reserve(npx * sizeof(float4)));
    ATX_HIP(r->comp.reserve(npx * sizeof(float4)));
    r->width = width;
    r->height = height;
    r->frame_index = 1;
    return ATX_OK;
}

int atx_renderer_set_scene(atx_renderer* h, const atx_sphere* spheres, uint64_t num_spheres,
                           const atx_material* materials, uint64_t num_materials, const atx_light* lights,
                           uint64_t num_lights)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r, "renderer is NULL");
    ATX_HIP(hipStreamSynchronize(r->stream));
    r->first_valid = r->prev_valid = false;
    int rc = upload_scene(spheres, num_spheres, materials, num_materials, lights, num_lights, &r->geo);
    if (rc != ATX_OK) return rc;
    r->has_scene = true;
    return ATX_OK;
}

int atx_renderer_set_kernel(atx_renderer* h, int kind)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r, "renderer is NULL");
    ATX_CHECK_ARG(kind == ATX_KERNEL_PLAIN || kind == ATX_KERNEL_WAVEFRONT, "unknown kernel kind %d", kind);
    r->kernel = kind;
    return ATX_OK;
}

int atx_renderer_set_accel(atx_renderer* h, int accel)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r, "renderer is NULL");
    ATX_CHECK_ARG(accel == ATX_ACCEL_BVH || accel == ATX_ACCEL_NONE, "unknown accel %d", accel);
    r->accel = accel;
    return ATX_OK;
}

int atx_renderer_set_settings(atx_renderer* h, const atx_settings* s)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r && s, "NULL argument");
    ATX_CHECK_ARG(s->max_bounces >= 0, "max_bounces must be >= 0");
    r->settings = *s;
    return ATX_OK;
}

int atx_renderer_get_settings(const atx_renderer* h, atx_settings* s)
{
    const Renderer* r = reinterpret_cast<const Renderer*>(h);
    ATX_CHECK_ARG(r && s, "NULL argument");
    *s = r->settings;
    return ATX_OK;
}

int atx_renderer_reserve(atx_renderer* h, uint32_t frames)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r, "renderer is NULL");
    ATX_CHECK_ARG(r->width > 0 && r->height > 0, "reserve before on_resize");
    if (frames == 0) return ATX_OK;
    const size_t npx = (size_t)r->width * r->height;
    // the largest launch render() forms for such a call (as it sizes them)
    const uint32_t fit = (uint32_t)std::max<size_t>(1, kSampleBufferBytes / (npx * kSampleBytes));
    const uint32_t per_launch = std::min(std::min(frames, r->wf_frames), fit);
    ATX_HIP(r->wf_samples.reserve(npx * per_launch * kSampleBytes));
    // growing a buffer drops its contents: the per-camera caches are redone
    if (r->wf_dirs.bytes() < npx * sizeof(float4)) {
        ATX_HIP(r->wf_dirs.reserve(npx * sizeof(float4)));
        r->dirs_valid = r->first_valid = false;
    }
    if (r->has_scene && r->wf_first.bytes() < first_cache_bytes(npx, r->geo.n_light)) {
        ATX_HIP(r->wf_first.reserve(first_cache_bytes(npx, r->geo.n_light)));
        r->first_valid = false;
    }
    return ATX_OK;
}

int atx_renderer_reset_frame_index(atx_renderer* h)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r, "renderer is NULL");
    r->frame_index = 1;
    return ATX_OK;
}

int atx_renderer_get_frame_index(const atx_renderer* h, uint32_t* fi)
{
    const Renderer* r = reinterpret_cast<const Renderer*>(h);
    ATX_CHECK_ARG(r && fi, "NULL argument");
    *fi = r->frame_index;
    return ATX_OK;
}

int atx_renderer_set_accumulator(atx_renderer* h, int kind)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r, "renderer is NULL");
    ATX_CHECK_ARG(kind == ATX_ACCUM_NAIVE || kind == ATX_ACCUM_KAHAN, "unknown accumulator kind %d", kind);
    if (kind != r->accumulator) r->frame_index = 1;  // the buffers' meaning changes
    r->accumulator = kind;
    return ATX_OK;
}

int atx_renderer_render(atx_renderer* h, const atx_camera* cam)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r, "renderer is NULL");
    return renderer_launch_frames(r, cam, 1, nullptr, nullptr, nullptr, 0);
}

int atx_renderer_render_frames(atx_renderer* h, const atx_camera* cam, uint32_t frames)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r, "renderer is NULL");
    return renderer_launch_frames(r, cam, frames, nullptr, nullptr, nullptr, 0);
}

int atx_renderer_render_debug(atx_renderer* h, const atx_camera* cam, float* radiance, int32_t* hit_ids,
                              int32_t* shadow_ids, uint32_t max_bounces)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r, "renderer is NULL");
    ATX_CHECK_ARG(!(hit_ids || shadow_ids) || max_bounces > 0, "max_bounces must be > 0 for hit capture");
    return renderer_launch_frames(r, cam, 1, radiance, hit_ids, shadow_ids, max_bounces);
}

int atx_renderer_get_image(atx_renderer* h, uint32_t* rgba, uint64_t n)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r && rgba, "NULL argument");
    const uint64_t npx = (uint64_t)r->width * r->height;
    ATX_CHECK_ARG(n >= npx, "buffer holds %llu pixels, image has %llu", (unsigned long long)n,
                  (unsigned long long)npx);
    if (npx == 0) return ATX_OK;
    ATX_HIP(hipMemcpyAsync(rgba, r->image.get(), npx * 4, hipMemcpyDeviceToHost, r->stream));
    ATX_HIP(hipStreamSynchronize(r->stream));
    return ATX_OK;
}

int atx_renderer_get_accumulation(atx_renderer* h, float* rgba, uint64_t n)
{
    Renderer* r = reinterpret_cast<Renderer*>(h);
    ATX_CHECK_ARG(r && rgba, "NULL argument");
    const uint64_t nf = 4ull * r->width * r->height;
    ATX_CHECK_ARG(n >= nf, "buffer holds %llu floats, need %llu", (unsigned long long)n, (unsigned long long)nf);
    if (nf == 0) return ATX_OK;
    ATX_HIP(hipMemcpyAsync(rgba, r->accum.get(), nf * 4, hipMemcpyDeviceToHost, r->stream));
    ATX_HIP(hipStreamSynchronize(r->stream));
    return ATX_OK;
}

int atx_renderer_get_stats(const atx_renderer* h, atx_render_stats* s)
{
    const Renderer* r = reinterpret_cast<const Renderer*>(h);
    ATX_CHECK_ARG(r && s, "NULL argument");
    *s = r->stats;
    return ATX_OK;
}

}  // extern "C"
