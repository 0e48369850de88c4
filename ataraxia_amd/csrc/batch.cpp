// batch.cpp -- many scenes of uneven size, sphere count, bounce depth and spp
// rendered by the persistent wavefront kernel.
//
// Launch L renders frames [L*F + 1, L*F + F] (F = frames per launch) of every
// scene whose spp exceeds L*F, each scene with min(F, spp - L*F) frames.
// Scenes are ordered by spp (descending), so the scenes of launch L are a
// prefix of the order.  Every launch has its own scene table (frame range
// and chunk-id range per scene) and its own chunk counter; all launches (plus
// a kernel that zeroes the counters) are captured once into a HIP graph and
// replayed by every render call.  Chunk size, frames per launch and grid size
// come from atx_batch_set_tuning or from the autotuner (autotune.cpp), whose
// choice is cached per device and workload.  Every scene of two or more
// frames builds its primary directions and first-bounce cache once, when the
// batch is first prepared; launches then run paths only for its hit pixels.
#include <algorithm>
#include <memory>
#include <new>
#include <numeric>
#include <vector>

#include "batch_internal.h"
#include "device_types.h"
#include "host_common.h"
#include "wavefront_host.h"

namespace atx {

struct BatchEntry {
    DevScene geo;
    DevBuf accum, comp, image;
    DevBuf dirs;                  // primary ray direction per pixel (written once)
    DevBuf first;                 // first-bounce cache per pixel (written once)
    uint32_t n_hit = 0;           // pixels of its hit list
    bool dirs_ready = false;
    atx_settings settings{};
    atx_camera cam{};
    uint32_t spp = 1;
    DevBuf dbg_rad, dbg_hit, dbg_sh;
    uint32_t dbg_bounces = 0;
    bool capture = false;
};

struct Batch {
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<std::unique_ptr<BatchEntry>> scenes;
    int accumulator = ATX_ACCUM_NAIVE;
    int accel = ATX_ACCEL_BVH;
    atx_batch_tuning tuning{0, 0, 1, 0, 0};
    DevBuf d_table, d_queue, counters;
    std::vector<uint32_t> order;          // table position -> scene id
    std::vector<uint32_t> launch_off;     // first table entry of launch l
    std::vector<uint32_t> launch_scenes;  // scenes taking part in launch l
    std::vector<uint32_t> launch_chunks;  // chunk ids of launch l
    std::vector<uint32_t> launch_chunk_px;  // jobs per chunk of launch l
    uint32_t launches = 0, chunk_px = 0, frames_per_launch = 0, grid = 0, max_spp = 0;
    uint64_t skipped = 0;   // primary rays of pixels left out of the jobs (hit lists), per render
    bool flat_only = false;   // no scene has a tree: the flat-loop kernel instance
    bool kahan = false;
    DevBuf d_samples;                     // sample buffer, shared by the launches (stream order)
    std::vector<uint32_t> launch_px;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;
    uint32_t prepared_frames = 0;
    std::vector<uint32_t> subset;          // autotuning: render only these scene ids (empty = all)
    atx_render_stats stats{};
};

static void drop_graph(Batch* b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

static int enqueue_launches(Batch* b)
{
    ATX_HIP(launch_zero_u32(b->d_queue.as<unsigned int>(), b->launches, b->stream));
    for (uint32_t l = 0; l < b->launches; ++l) {
        WfPass wp{};
        wp.scenes = b->d_table.as<WfScene>() + b->launch_off[l];
        wp.n_scenes = b->launch_scenes[l];
        wp.total_chunks = b->launch_chunks[l];
        wp.chunk_px = b->launch_chunk_px[l];
        wp.total_px = b->launch_px[l];
        wp.queue = b->d_queue.as<unsigned int>() + l;
        wp.counters = b->counters.as<unsigned long long>();
        ATX_HIP(launch_wavefront(wp, WfLaunch{b->grid, b->flat_only, b->tuning.waves_per_simd}, b->stream));
    }
    return ATX_OK;
}

// Builds the per-launch scene tables; captures the graph.  max_frames limits
// the rendered frames (autotuning renders only the first launch).
static int batch_prepare(Batch* b, uint32_t max_frames)
{
    drop_graph(b);
    if (b->subset.empty()) {
        b->order.resize(b->scenes.size());
        std::iota(b->order.begin(), b->order.end(), 0u);
    } else {
        b->order = b->subset;
    }
    const size_t n = b->order.size();
    std::stable_sort(b->order.begin(), b->order.end(),
                     [b](uint32_t x, uint32_t y) { return b->scenes[x]->spp > b->scenes[y]->spp; });
    b->chunk_px = b->tuning.chunk_px;   // 0: auto_chunk_px per launch
    b->frames_per_launch = b->tuning.frames_per_launch ? b->tuning.frames_per_launch : kDefaultFramesPerLaunch;
    b->kahan = b->accumulator == ATX_ACCUM_KAHAN;
    b->flat_only = batch_flat_only(b);
    b->grid = b->tuning.workgroups ? b->tuning.workgroups : wavefront_full_grid(b->flat_only, b->tuning.waves_per_simd);
    b->max_spp = 0;
    for (auto& e : b->scenes) b->max_spp = std::max(b->max_spp, std::min(e->spp, max_frames));
    const uint32_t F = b->frames_per_launch;
    std::vector<WfScene> table;
    b->launch_off.clear();
    b->launch_scenes.clear();
    b->launch_chunks.clear();
    b->launch_chunk_px.clear();
    b->launch_px.clear();
    // Launch (f0, group) renders frames [f0 + 1, f0 + F] of one group of the
    // scenes whose spp exceeds f0: consecutive scenes in spp order whose
    // samples fit kSampleBufferBytes (at least one scene per group).
    for (uint32_t f0 = 0; f0 < b->max_spp; f0 += F) {
        b->launch_off.push_back((uint32_t)table.size());
        uint32_t k = 0;
        size_t group_bytes = 0;
        for (; k < n; ++k) {
            BatchEntry& e = *b->scenes[b->order[k]];
            const uint32_t spp = std::min(e.spp, max_frames);
            if (spp <= f0) break;
            const size_t bytes = (size_t)e.cam.width * e.cam.height * std::min(F, spp - f0) * kSampleBytes;
            if (group_bytes > 0 && group_bytes + bytes > kSampleBufferBytes) {
                b->launch_scenes.push_back((uint32_t)(table.size() - b->launch_off.back()));
                b->launch_off.push_back((uint32_t)table.size());
                group_bytes = 0;
            }
            group_bytes += bytes;
            WfScene w;
            fill_wf_scene(e.geo, e.settings, e.cam, e.accum.as<float4>(),
                          b->kahan ? e.comp.as<float4>() : nullptr, e.image.as<uint32_t>(), b->kahan,
                          f0 + 1u, std::min(F, spp - f0), b->accel, &w);
            // the first-bounce cache pays from two frames per scene on
            const bool cache = e.spp >= 2;
            const size_t npx = (size_t)e.cam.width * e.cam.height;
            if (!e.dirs_ready) {
                ATX_HIP(e.dirs.reserve(npx * sizeof(float4)));
                if (cache) ATX_HIP(e.first.reserve(first_cache_bytes(npx, w.n_light)));
            }
            w.dirs = e.dirs.as<float4>();
            if (cache) set_first_cache(w, e.first.get());
            if (!e.dirs_ready) {
                ATX_HIP(launch_camera_dirs(w, e.dirs.as<float4>(), b->stream));
                if (cache) {
                    ATX_HIP(launch_first_hits(w, b->stream));
                    ATX_HIP(read_first_hits(e.first.get(), npx, w.n_light, b->stream, &e.n_hit));
                }
                e.dirs_ready = true;
            }
            if (e.capture) {
                w.dbg_radiance = e.dbg_rad.as<float4>();
                w.dbg_hit = e.dbg_hit.as<int32_t>();
                w.dbg_shadow = e.dbg_sh.as<int32_t>();
                w.dbg_bounces = e.dbg_bounces;
            }
            if (cache) use_hit_list(w, e.first.get(), e.n_hit);
            table.push_back(w);
        }
        b->launch_scenes.push_back((uint32_t)(table.size() - b->launch_off.back()));
    }
    // the launches run one after another on the stream: one sample buffer serves all
    size_t most = 0;
    for (size_t l = 0; l < b->launch_off.size(); ++l)
        most = std::max(most, launch_samples(table.data() + b->launch_off[l], b->launch_scenes[l]));
    ATX_HIP(b->d_samples.reserve(std::max<size_t>(1, most) * kSampleBytes));
    for (size_t l = 0; l < b->launch_off.size(); ++l) {
        const WfScene* lt = table.data() + b->launch_off[l];
        const uint32_t P = b->chunk_px ? b->chunk_px : auto_chunk_px(launch_jobs(lt, b->launch_scenes[l]), b->grid);
        WfPass lp{};
        if (!layout_launch(table.data() + b->launch_off[l], b->launch_scenes[l], P,
                           b->d_samples.as<float3>(), &lp)) {
            set_error("batch: a launch of %u frames holds 2^31 or more paths; lower frames_per_launch", F);
            return ATX_ERR_INVALID;
        }
        b->launch_chunks.push_back(lp.total_chunks);
        b->launch_chunk_px.push_back(P);
        b->launch_px.push_back(lp.total_px);
    }
    b->launches = (uint32_t)b->launch_off.size();
    b->skipped = skipped_primaries(table.data(), table.size());
    ATX_HIP(b->d_table.reserve(std::max<size_t>(1, table.size()) * sizeof(WfScene)));
    if (!table.empty())
        ATX_HIP(hipMemcpy(b->d_table.get(), table.data(), table.size() * sizeof(WfScene), hipMemcpyHostToDevice));
    ATX_HIP(b->d_queue.reserve(std::max<uint32_t>(1, b->launches) * sizeof(unsigned int)));
    if (b->tuning.use_graph && b->launches) {
        ATX_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_launches(b);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc != ATX_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        ATX_HIP(e);
        b->graph = g;
        ATX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    b->prepared_frames = max_frames;
    b->dirty = false;
    return ATX_OK;
}

int batch_run(Batch* b, uint32_t max_frames, double* ms_out)
{
    if (b->scenes.empty()) {
        set_error("batch has no scenes");
        return ATX_ERR_STATE;
    }
    if (b->dirty || b->prepared_frames != max_frames) {
        int rc = batch_prepare(b, max_frames);
        if (rc != ATX_OK) return rc;
    }
    ATX_HIP(hipMemsetAsync(b->counters.get(), 0, kCtrCount * sizeof(unsigned long long), b->stream));
    ATX_HIP(hipEventRecord(b->ev0, b->stream));
    if (b->exec) {
        ATX_HIP(hipGraphLaunch(b->exec, b->stream));
    } else {
        int rc = enqueue_launches(b);
        if (rc != ATX_OK) return rc;
    }
    ATX_HIP(hipEventRecord(b->ev1, b->stream));
    unsigned long long ctr[kCtrCount];
    ATX_HIP(hipMemcpyAsync(ctr, b->counters.get(), sizeof(ctr), hipMemcpyDeviceToHost, b->stream));
    ATX_HIP(hipStreamSynchronize(b->stream));
    float ms = 0.0f;
    ATX_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    uint64_t samples = 0;
    for (uint32_t id : b->order) {
        const BatchEntry& e = *b->scenes[id];
        const uint64_t frames = std::min<uint64_t>(e.spp, b->prepared_frames);
        samples += (uint64_t)e.cam.width * e.cam.height * frames;
    }
    b->stats.samples = samples;
    b->stats.rays = ctr[kCtrRays] + ctr[kCtrReused] + b->skipped;
    b->stats.reused_rays = ctr[kCtrReused] + b->skipped;
    b->stats.sphere_tests = ctr[kCtrTests];
    b->stats.exact_retraces = ctr[kCtrExact];
    b->stats.box_tests = ctr[kCtrBoxes];
    b->stats.gpu_ms = ms;
    b->stats.launches = b->launches;
    b->stats.path_ms = 0.0;
    if (ms_out) *ms_out = ms;
    return ATX_OK;
}

Batch* batch_cast(atx_batch* h) { return reinterpret_cast<Batch*>(h); }

bool batch_is_kahan(const Batch* b) { return b->accumulator == ATX_ACCUM_KAHAN; }

int batch_accel(const Batch* b) { return b->accel; }

bool batch_flat_only(const Batch* b)
{
    for (const auto& e : b->scenes)
        if (e->geo.accel(b->accel).n_nodes != 0)
            return false;
    return true;
}

void batch_set_subset(Batch* b, const std::vector<uint32_t>& ids)
{
    b->subset = ids;
    b->dirty = true;
}

size_t batch_size(const Batch* b) { return b->scenes.size(); }

uint64_t batch_scene_pixels(const Batch* b, uint32_t id)
{
    const BatchEntry& e = *b->scenes[id];
    return (uint64_t)e.cam.width * e.cam.height;
}

void batch_signature(const Batch* b, uint64_t* scenes, uint64_t* pixels, uint64_t* spheres, uint64_t* spp,
                     uint64_t* bounces)
{
    *scenes = b->scenes.size();
    *pixels = *spheres = *spp = *bounces = 0;
    for (auto& e : b->scenes) {
        *pixels += (uint64_t)e->cam.width * e->cam.height;
        *spheres += e->geo.n_sph;
        *spp += e->spp;
        *bounces += (uint64_t)e->settings.max_bounces;
    }
}

static BatchEntry* entry(Batch* b, uint32_t id)
{
    if (!b || id >= b->scenes.size()) {
        set_error("unknown scene id %u", id);
        return nullptr;
    }
    return b->scenes[id].get();
}

}  // namespace atx

using namespace atx;

extern "C" {

int atx_batch_create(atx_batch** out)
{
    ATX_CHECK_ARG(out, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        set_error("no HIP device available");
        return ATX_ERR_NODEVICE;
    }
    Batch* b = new (std::nothrow) Batch();
    if (!b) { set_error("out of host memory"); return ATX_ERR_NOMEM; }
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = b->counters.reserve(kCtrCount * sizeof(unsigned long long));
    if (e != hipSuccess) {
        set_error("batch create: %s", hipGetErrorString(e));
        atx_batch_destroy(reinterpret_cast<atx_batch*>(b));
        return ATX_ERR_HIP;
    }
    *out = reinterpret_cast<atx_batch*>(b);
    return ATX_OK;
}

int atx_batch_destroy(atx_batch* h)
{
    Batch* b = batch_cast(h);
    if (!b) return ATX_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return ATX_OK;
}

int atx_batch_add_scene(atx_batch* h, const atx_scene_desc* d, uint32_t* scene_id)
{
    Batch* b = batch_cast(h);
    ATX_CHECK_ARG(b && d, "NULL argument");
    ATX_CHECK_ARG(d->camera.width > 0 && d->camera.height > 0, "Width or height cannot be zero");
    ATX_CHECK_ARG((uint64_t)d->camera.width * d->camera.height < (1ull << 31), "viewport too large");
    ATX_CHECK_ARG(d->settings.max_bounces >= 0, "max_bounces must be >= 0");
    ATX_CHECK_ARG(d->spp >= 1, "spp must be >= 1");
    std::unique_ptr<BatchEntry> e(new (std::nothrow) BatchEntry());
    if (!e) { set_error("out of host memory"); return ATX_ERR_NOMEM; }
    int rc = upload_scene(d->spheres, d->num_spheres, d->materials, d->num_materials, d->lights, d->num_lights,
                          &e->geo);
    if (rc != ATX_OK) return rc;
    const size_t npx = (size_t)d->camera.width * d->camera.height;
    ATX_HIP(e->accum.reserve(npx * sizeof(float4)));
    ATX_HIP(e->comp.reserve(npx * sizeof(float4)));
    ATX_HIP(e->image.reserve(npx * sizeof(uint32_t)));
    e->settings = d->settings;
    e->cam = d->camera;
    e->spp = d->spp;
    if (scene_id) *scene_id = (uint32_t)b->scenes.size();
    b->scenes.push_back(std::move(e));
    b->dirty = true;
    return ATX_OK;
}

int atx_batch_set_accumulator(atx_batch* h, int kind)
{
    Batch* b = batch_cast(h);
    ATX_CHECK_ARG(b, "batch is NULL");
    ATX_CHECK_ARG(kind == ATX_ACCUM_NAIVE || kind == ATX_ACCUM_KAHAN, "unknown accumulator kind %d", kind);
    if (kind != b->accumulator) b->dirty = true;
    b->accumulator = kind;
    return ATX_OK;
}

int atx_batch_set_accel(atx_batch* h, int accel)
{
    Batch* b = batch_cast(h);
    ATX_CHECK_ARG(b, "batch is NULL");
    ATX_CHECK_ARG(accel == ATX_ACCEL_BVH || accel == ATX_ACCEL_NONE, "unknown accel %d", accel);
    if (accel != b->accel) b->dirty = true;
    b->accel = accel;
    return ATX_OK;
}

int atx_batch_set_tuning(atx_batch* h, const atx_batch_tuning* t)
{
    Batch* b = batch_cast(h);
    ATX_CHECK_ARG(b && t, "NULL argument");
    ATX_CHECK_ARG(t->chunk_px == 0 || (t->chunk_px % 64 == 0 && t->chunk_px <= kMaxChunkPx),
                  "chunk_px must be a multiple of 64, at most 1024");
    ATX_CHECK_ARG(t->frames_per_launch <= 4096, "frames_per_launch must be <= 4096");
    ATX_CHECK_ARG(t->workgroups <= 65536, "workgroups too large");
    ATX_CHECK_ARG(t->waves_per_simd == 0 || (t->waves_per_simd >= kWfWavesMin && t->waves_per_simd <= kWfWavesMax),
                  "waves_per_simd must be 0 or %u..%u", kWfWavesMin, kWfWavesMax);
    b->tuning = *t;
    b->dirty = true;
    return ATX_OK;
}

int atx_batch_get_tuning(const atx_batch* h, atx_batch_tuning* t)
{
    const Batch* b = reinterpret_cast<const Batch*>(h);
    ATX_CHECK_ARG(b && t, "NULL argument");
    *t = b->tuning;
    if (!t->frames_per_launch) t->frames_per_launch = kDefaultFramesPerLaunch;
    const bool flat = batch_flat_only(b);
    t->waves_per_simd = wf_waves(t->waves_per_simd, flat);
    if (!t->workgroups)
        t->workgroups = b->grid ? b->grid : wavefront_full_grid(flat, t->waves_per_simd);
    return ATX_OK;
}

int atx_batch_capture(atx_batch* h, uint32_t scene_id, uint32_t max_bounces)
{
    Batch* b = batch_cast(h);
    BatchEntry* e = entry(b, scene_id);
    if (!e) return ATX_ERR_NOTFOUND;
    ATX_CHECK_ARG(max_bounces > 0 && max_bounces <= 1024, "max_bounces must be in [1, 1024]");
    const size_t npx = (size_t)e->cam.width * e->cam.height;
    ATX_HIP(e->dbg_rad.reserve(npx * sizeof(float4)));
    ATX_HIP(e->dbg_hit.reserve(npx * max_bounces * 4));
    ATX_HIP(e->dbg_sh.reserve(npx * max_bounces * 4));
    e->dbg_bounces = max_bounces;
    e->capture = true;
    b->dirty = true;
    return ATX_OK;
}

int atx_batch_get_capture(atx_batch* h, uint32_t scene_id, float* radiance, int32_t* hit_ids,
                          int32_t* shadow_ids)
{
    Batch* b = batch_cast(h);
    BatchEntry* e = entry(b, scene_id);
    if (!e) return ATX_ERR_NOTFOUND;
    if (!e->capture) { set_error("capture not enabled for scene %u", scene_id); return ATX_ERR_STATE; }
    const size_t npx = (size_t)e->cam.width * e->cam.height;
    if (radiance) ATX_HIP(hipMemcpy(radiance, e->dbg_rad.get(), npx * sizeof(float4), hipMemcpyDeviceToHost));
    if (hit_ids) ATX_HIP(hipMemcpy(hit_ids, e->dbg_hit.get(), npx * e->dbg_bounces * 4, hipMemcpyDeviceToHost));
    if (shadow_ids) ATX_HIP(hipMemcpy(shadow_ids, e->dbg_sh.get(), npx * e->dbg_bounces * 4, hipMemcpyDeviceToHost));
    return ATX_OK;
}

int atx_batch_render(atx_batch* h)
{
    Batch* b = batch_cast(h);
    ATX_CHECK_ARG(b, "batch is NULL");
    return batch_run(b, 0xFFFFFFFFu, nullptr);
}

int atx_batch_render_frames(atx_batch* h, uint32_t max_frames)
{
    Batch* b = batch_cast(h);
    ATX_CHECK_ARG(b, "batch is NULL");
    ATX_CHECK_ARG(max_frames >= 1, "max_frames must be >= 1");
    return batch_run(b, max_frames, nullptr);
}

int atx_batch_get_image(atx_batch* h, uint32_t scene_id, uint32_t* rgba, uint64_t n)
{
    Batch* b = batch_cast(h);
    BatchEntry* e = entry(b, scene_id);
    if (!e) return ATX_ERR_NOTFOUND;
    const uint64_t npx = (uint64_t)e->cam.width * e->cam.height;
    ATX_CHECK_ARG(rgba && n >= npx, "buffer too small (%llu < %llu)", (unsigned long long)n, (unsigned long long)npx);
    ATX_HIP(hipMemcpy(rgba, e->image.get(), npx * 4, hipMemcpyDeviceToHost));
    return ATX_OK;
}

int atx_batch_get_accumulation(atx_batch* h, uint32_t scene_id, float* rgba, uint64_t n)
{
    Batch* b = batch_cast(h);
    BatchEntry* e = entry(b, scene_id);
    if (!e) return ATX_ERR_NOTFOUND;
    const uint64_t nf = 4ull * e->cam.width * e->cam.height;
    ATX_CHECK_ARG(rgba && n >= nf, "buffer too small (%llu < %llu)", (unsigned long long)n, (unsigned long long)nf);
    ATX_HIP(hipMemcpy(rgba, e->accum.get(), nf * 4, hipMemcpyDeviceToHost));
    return ATX_OK;
}

int atx_batch_get_stats(const atx_batch* h, atx_render_stats* s)
{
    const Batch* b = reinterpret_cast<const Batch*>(h);
    ATX_CHECK_ARG(b && s, "NULL argument");
    *s = b->stats;
    return ATX_OK;
}

}  // extern "C"
