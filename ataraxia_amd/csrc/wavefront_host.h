// wavefront_host.h -- host helpers that fill the persistent kernel's scene
// table entries (shared by the batch renderer, the server and the Renderer's
// wavefront mode).
#pragma once

#include <string.h>

#include "camera_host.h"
#include "host_common.h"
#include "wavefront.h"

namespace atx {

inline void fill_wf_scene(const DevScene& g, const atx_settings& st, const atx_camera& cam, float4* accum,
                          float4* comp, uint32_t* image, bool kahan, uint32_t frame0, uint32_t frames,
                          int accel, WfScene* out)
{
    memset(out, 0, sizeof(*out));
    out->sph = g.sph.as<float4>();
    out->sph_mat = g.sph_mat.as<int32_t>();
    out->mat = g.mat.as<float4>();
    out->light = g.light.as<float4>();
    out->n_sph = g.n_sph;
    out->n_mat = g.n_mat;
    out->n_light = g.n_light;
    out->width = cam.width;
    out->height = cam.height;
    out->max_bounces = st.max_bounces;
    out->sky_light = st.sky_light ? 1u : 0u;
    out->frame0 = frame0;
    out->frames = frames;
    out->kahan = kahan ? 1u : 0u;
    camera_inverse_matrices(cam.position, cam.direction, cam.fov_deg, cam.near_clip, cam.far_clip, cam.width,
                            cam.height, out->inv_proj, out->inv_view);
    out->cam_pos[0] = cam.position[0];
    out->cam_pos[1] = cam.position[1];
    out->cam_pos[2] = cam.position[2];
    out->accum = accum;
    out->comp = comp;
    out->image = image;
    out->bvh = g.accel(accel);
}


// Samples (kSampleBytes each) one launch of these table entries needs.
inline size_t launch_samples(const WfScene* tab, size_t n)
{
    size_t s = 0;
    for (size_t k = 0; k < n; ++k)
        s += (size_t)tab[k].width * tab[k].height * tab[k].frames;
    return s;
}

// Paths (jobs) of one table entry: every pixel's, or only those of the hit
// list (hit_px) in every frame.
inline uint64_t scene_jobs(const WfScene& w)
{
    return (w.hit_px ? (uint64_t)w.n_hit : (uint64_t)w.width * w.height) * w.frames;
}
inline uint64_t launch_jobs(const WfScene* tab, size_t n)
{
    uint64_t s = 0;
    for (size_t k = 0; k < n; ++k)
        s += scene_jobs(tab[k]);
    return s;
}

// Reads the hit count k_first_hit left in buf (a first-bounce cache of
// `pixels` pixels and `lights` lights); waits for `stream`.
inline hipError_t read_first_hits(void* buf, size_t pixels, uint32_t lights, hipStream_t stream, uint32_t* n)
{
    const FirstCacheView v = first_cache_view(buf, pixels, lights);
    hipError_t e = hipMemcpyAsync(n, v.n_hits, sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    return e;
}

// Primary rays of the pixels a table entry leaves out of its jobs (hit
// list): the reference traces them (they miss), so statistics count them as
// rays answered from the cache.
inline uint64_t skipped_primaries(const WfScene* tab, size_t n)
{
    uint64_t s = 0;
    for (size_t k = 0; k < n; ++k)
        if (tab[k].hit_px)
            s += ((uint64_t)tab[k].width * tab[k].height - tab[k].n_hit) * tab[k].frames;
    return s;
}

// Restricts a table entry's jobs to the cache's hit list: only with a
// first-bounce cache, paths of at least one bounce, and no debug capture
// (the capture records every pixel).
inline void use_hit_list(WfScene& w, void* buf, uint32_t n_hit)
{
    if (!w.first || w.max_bounces <= 0 || w.dbg_radiance || w.dbg_hit || w.dbg_shadow)
        return;
    w.hit_px = first_cache_view(buf, (size_t)w.width * w.height, w.n_light).hits;
    w.n_hit = n_hit;
}

// Gives the launch's table entries their chunk ids (ceil(pixels * frames /
// chunk_px) chunks of chunk_px jobs each), pixel offsets and slices of the
// sample buffer (in table order) and fills the launch's totals.  False when
// a count would not fit the kernel's 32-bit indices (a scene's jobs, or the
// launch's pixels or chunks, reach 2^31).
inline bool layout_launch(WfScene* tab, size_t n, uint32_t chunk_px, float3* samples, WfPass* p)
{
    const uint64_t kLimit = 1ull << 31;
    uint64_t chunks = 0, px = 0;
    size_t off = 0;
    for (size_t k = 0; k < n; ++k) {
        WfScene& w = tab[k];
        const uint64_t npx = (uint64_t)w.width * w.height;
        const uint64_t jobs = scene_jobs(w);
        const uint64_t nch = (jobs + chunk_px - 1) / chunk_px;
        if (jobs >= kLimit || chunks + nch >= kLimit || px + npx >= kLimit)
            return false;
        w.chunk_begin = (uint32_t)chunks;
        w.n_chunks = (uint32_t)nch;
        w.px_begin = (uint32_t)px;
        w.samples = samples + off;
        w.div_group = udiv_make(kJobGroupPx * w.frames);
        w.div_light = udiv_make(w.n_light ? w.n_light : 1u);
        chunks += nch;
        px += npx;
        off += (size_t)(npx * w.frames);
    }
    p->total_chunks = (uint32_t)chunks;
    p->total_px = (uint32_t)px;
    p->chunk_px = chunk_px;
    return true;
}

inline uint32_t device_cus()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    return (uint32_t)cus;
}

// Persistent grid that fills the device: CUs x resident workgroups per CU.
inline uint32_t wavefront_full_grid(bool flat_only, uint32_t waves = 0)
{
    return device_cus() * (uint32_t)wavefront_max_blocks_per_cu(flat_only, waves);
}

// True when no scene of the table has a tree (the flat-loop kernel instance).
inline bool wf_flat_only(const WfScene* tab, size_t n)
{
    for (size_t k = 0; k < n; ++k)
        if (tab[k].bvh.n_nodes != 0)
            return false;
    return true;
}

// Jobs per chunk for a launch of `jobs` jobs on a persistent grid of
// `workgroups` (4 waves each) when the caller gives none.  Small chunks share
// the launch out evenly to the end; large ones keep a wave on neighbouring
// pixels (and in a many-frame launch on the same pixel) and fetch fewer
// blocks.  Measured on the 1080p workload (1 to 64 frames per launch, chunks
// of 64 to 1024 jobs, both kernel instances): 512 jobs from about 5000 jobs
// per wave on (16 frames and more), 256 below; within 6 % of the best chunk
// at every frame count.
inline uint32_t auto_chunk_px(uint64_t jobs, uint32_t workgroups)
{
    const uint64_t per_wave = jobs / (4ull * (workgroups ? workgroups : 1u));
    return per_wave >= 5000u ? 512u : 256u;
}

// Frames per launch of batches and serving ticks (each scene or request
// advances this many frames per launch).
constexpr uint32_t kDefaultFramesPerLaunch = 64;
// Frames per launch of the single-scene renderer.  Every launch of the
// persistent kernel costs about 0.7 ms beyond its paths (ramp-up and
// drain; 1080p: 10.75 ms for 64 frames, 5.73 ms for 32), so a render call
// takes as few launches as its sample buffer allows: 1080p, 256 frames per
// launch, 0.172 -> 0.160 ms per frame against 64.
constexpr uint32_t kRendererFramesPerLaunch = 256;
// Upper bound of one launch's sample buffer (kSampleBytes per pixel and frame):
// renderers shorten their launches, batches split them by scene groups.
// 8 GiB is 3 % of the device's HBM and holds 345 frames of 1080p.
constexpr size_t kSampleBufferBytes = size_t(8) << 30;

}  // namespace atx
