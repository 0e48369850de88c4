// wavefront.h -- argument blocks of the persistent (path-regenerating) kernel
// shared by the batched renderer, the serving loop and the Renderer's
// wavefront mode.
//
// One launch renders frames [frame0, frame0 + frames) of every scene in its
// table in two kernels.  The path kernel treats every (frame, pixel) path as
// an independent job; a scene's jobs come in groups of kJobGroupPx
// neighbouring pixels times the launch's frames (job_pixel_frame), over all
// of its pixels or only those of its hit list.  A wave's first block is its
// grid index, later ones come from a global counter (scene s owns chunk ids
// [chunk_begin, chunk_begin + n_chunks), n_chunks = ceil(jobs / chunk_px));
// it hands the block's paths to its lanes one by one; a lane whose path ends
// stores its sample in the launch's sample buffer and takes the next job, so every
// closest-hit loop runs with (nearly) all 64 lanes busy and no wave waits for
// another's work.  The accumulate kernel then adds each pixel's samples to
// its accumulation strictly in frame order and writes the RGBA8 value, which
// makes the result bit-identical to rendering one frame per launch.  A wave
// works on one scene at a time (spheres are read with wave-uniform scalar
// loads) and switches scenes once its lanes drain.
#pragma once

#include <stdint.h>
#include <hip/hip_runtime.h>

#include "bvh.h"

namespace atx {

// Division of 32-bit unsigned integers by a divisor that stays fixed for
// many divisions, by a multiply and two shifts (Granlund & Montgomery,
// "Division by Invariant Integers using Multiplication", PLDI 1994, Fig. 4.1):
// with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
//   t = mulhi(m, n),  floor(n / d) = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0)
// for all 0 <= n < 2^32 and 1 <= d < 2^32.  The path kernel divides by the
// scene's light count and job-group length once or more per path; the
// compiler's general sequence costs about 20 instructions, four of them
// quarter-rate multiplies.
struct UDiv {
    uint32_t d, m, s1, s2;
};
__host__ __device__ inline UDiv udiv_make(uint32_t d)
{
    uint32_t l = 0;
    while (l < 32u && (1ull << l) < d)
        ++l;
    UDiv r;
    r.d = d;
    r.m = (uint32_t)((((1ull << l) - d) << 32) / d + 1u);
    r.s1 = l < 1u ? l : 1u;
    r.s2 = l > 1u ? l - 1u : 0u;
    return r;
}
__host__ __device__ __forceinline__ uint32_t udiv(uint32_t n, const UDiv& v)
{
    const uint32_t t = (uint32_t)(((uint64_t)v.m * n) >> 32);
    return (t + ((n - t) >> v.s1)) >> v.s2;
}
__host__ __device__ __forceinline__ uint32_t umod(uint32_t n, const UDiv& v)
{
    return n - udiv(n, v) * v.d;
}

struct WfScene {
    // geometry (device layout of device_types.h)
    const float4* sph;
    const int32_t* sph_mat;
    const float4* mat;
    const float4* light;
    uint32_t n_sph;
    uint32_t n_mat;
    uint32_t n_light;
    uint32_t width;
    uint32_t height;
    int32_t max_bounces;
    uint32_t sky_light;
    uint32_t frame0;       // first frame index rendered by this launch (>= 1)
    uint32_t frames;       // frames rendered by this launch (>= 1)
    uint32_t chunk_begin;  // first chunk id of this scene in the launch
    uint32_t n_chunks;     // ceil(width * height * frames / chunk_px)
    uint32_t kahan;        // 1: compensated accumulation (comp must be valid)
    uint32_t dbg_bounces;  // stride of the debug arrays
    uint32_t px_begin;     // first pixel of this scene in the launch's pixel range
    float inv_proj[16];
    float inv_view[16];
    float cam_pos[4];
    float4* accum;
    float4* comp;
    uint32_t* image;
    float4* dbg_radiance;  // optional per-pixel capture of the last frame (nullptr = off)
    int32_t* dbg_hit;
    int32_t* dbg_shadow;
    float3* samples;       // this launch's samples (r, g, b), indexed by the scene's job
                           // (job_pixel_frame order); k_accumulate adds the alpha of 1 itself
    const float4* dirs;    // primary ray direction per pixel (x, y, z, 0); nullptr: computed per path
    const uint4* first;        // first-bounce records, 4 per pixel (launch_first_hits); nullptr: traced per path
    const float4* first_light; // first-bounce light terms, [light][pixel] (nullptr: none)
    // With a first-bounce cache, the pixels whose primary ray hits a sphere:
    // the launch's jobs cover only these, and k_accumulate adds the others'
    // sample, the sky colour (or black) of every frame, without a path.
    // nullptr: every pixel is a job.  Not with the debug capture or
    // max_bounces <= 0.  The kernels read the list's length from the cache
    // (job_slots), where k_first_hit leaves it, so a launch can follow
    // k_first_hit on the stream without the host reading the count first;
    // n_hit is the length the host lays the launch's chunks out for, at
    // least the list's (empty chunks past the list end the hand-out).
    const uint32_t* hit_px;
    uint32_t n_hit;
    BvhDev bvh;            // n_nodes = 0: flat loop over all spheres
    UDiv div_group;        // by kJobGroupPx * frames (set by layout_launch)
    UDiv div_light;        // by max(n_light, 1) (set by layout_launch)
};

struct WfPass {
    const WfScene* scenes;
    uint32_t n_scenes;
    uint32_t total_chunks; // chunk ids [0, total_chunks)
    uint32_t chunk_px;     // jobs per chunk (multiple of 64, <= kMaxChunkPx)
    uint32_t total_px;     // pixels of all scenes of the launch
    unsigned int* queue;   // this launch's chunk counter (zero before launch)
    unsigned long long* counters;
    uint32_t refill_min;   // idle lanes a wave waits for before it hands out jobs (set by launch_wavefront)
    uint32_t pad_;
};

// Idle lanes a wave of the path kernel collects before it hands out new
// jobs (unless every lane is idle); ATX_WF_REFILL (1-64) overrides it.
constexpr uint32_t kRefillMin = 16;

constexpr uint32_t kMaxChunkPx = 1024;

// Bytes of one path sample in the launch's sample buffer (r, g, b).
constexpr size_t kSampleBytes = 12;

// Job order.  A scene's jobs come in groups of kJobGroupPx neighbouring
// pixels (the last group may hold fewer) times the launch's frames, frame
// after frame inside a group: jobs 8g .. 8g+7 are pixels 8g .. 8g+7 of the
// first frame, the next 8 the same pixels in the second frame, and so on.
// The 64 lanes of a wave then run 8 frames of 8 neighbouring pixels at a
// time.  Samples are stored in job order, so a wave's block of chunk_px
// consecutive jobs (chunk_px * 12 bytes, a multiple of 128) fills whole
// 128-byte lines of the sample buffer that no other wave writes.
// "Pixel" here is a slot: a pixel, or an index into the scene's hit list.
constexpr uint32_t kJobGroupPx = 8;
__host__ __device__ __forceinline__ void job_pixel_frame(uint32_t j, uint32_t pixels, uint32_t frames, uint32_t& px,
                                                         uint32_t& frame)
{
    const uint32_t g = j / (kJobGroupPx * frames);
    const uint32_t w = j - g * kJobGroupPx * frames;
    const uint32_t left = pixels - kJobGroupPx * g;
    const uint32_t m = left < kJobGroupPx ? left : kJobGroupPx;
    frame = w / m;
    px = kJobGroupPx * g + (w - frame * m);
}
// The same with div_group = udiv_make(kJobGroupPx * frames): one multiply
// for the group, a shift for the frame but in a scene's last group.
__host__ __device__ __forceinline__ void job_pixel_frame(uint32_t j, uint32_t pixels, const UDiv& div_group,
                                                         uint32_t& px, uint32_t& frame)
{
    const uint32_t g = udiv(j, div_group);
    const uint32_t w = j - g * div_group.d;
    const uint32_t left = pixels - kJobGroupPx * g;
    const uint32_t m = left < kJobGroupPx ? left : kJobGroupPx;
    frame = m == kJobGroupPx ? w / kJobGroupPx : w / m;
    px = kJobGroupPx * g + (w - frame * m);
}

// The job of slot k in the launch's frame f (the inverse of job_pixel_frame).
__host__ __device__ __forceinline__ uint32_t job_of(uint32_t k, uint32_t f, uint32_t pixels, uint32_t frames)
{
    const uint32_t g = k / kJobGroupPx;
    const uint32_t left = pixels - kJobGroupPx * g;
    const uint32_t m = left < kJobGroupPx ? left : kJobGroupPx;
    return g * kJobGroupPx * frames + f * m + (k - g * kJobGroupPx);
}

// Resident waves per SIMD the path kernel's registers are budgeted for:
// launches with trees, and launches whose scenes all use the flat loop.
constexpr uint32_t kWfWavesTree = 6;
constexpr uint32_t kWfWavesFlat = 7;

// Register budgets the path kernel is compiled for (resident waves per
// SIMD); a launch picks one (the batch autotuner tries them all).
constexpr uint32_t kWfWavesMin = 5, kWfWavesMax = 8;
// Launches with trees of at least this many jobs take the 8-wave instance,
// which parks the path state in LDS during traces (1080p workload, 1.37 M
// jobs per frame, path kernel against 6 waves: 4 frames 4.5 % slower, 8
// frames even, 12 / 16 / 20 frames 0.7 / 6.3 / 8.2 % faster, 256 frames 6.8 %
// faster).
constexpr uint64_t kWfWaves8MinJobs = uint64_t(12) << 20;
__host__ __device__ __forceinline__ uint32_t wf_tree_waves(uint64_t jobs)
{
    return jobs >= kWfWaves8MinJobs ? 8u : kWfWavesTree;
}
// The instance a launch uses: `waves` when it names a compiled instance,
// else the default for its scenes.
__host__ __device__ __forceinline__ uint32_t wf_waves(uint32_t waves, bool flat_only)
{
    return waves >= kWfWavesMin && waves <= kWfWavesMax ? waves : (flat_only ? kWfWavesFlat : kWfWavesTree);
}

struct WfLaunch {
    uint32_t workgroups;     // persistent grid size (256-lane workgroups)
    bool flat_only = false;  // no scene of the launch has a tree
    uint32_t waves = 0;      // kernel instance (resident waves per SIMD); 0 = default
};

// The launch's path kernel, then its accumulate kernel; `paths_done`, when
// given, is recorded on the stream between the two.
hipError_t launch_wavefront(const WfPass& p, const WfLaunch& l, hipStream_t stream,
                            hipEvent_t paths_done = nullptr);
hipError_t launch_zero_u32(unsigned int* p, uint32_t n, hipStream_t stream);
// Writes the primary ray direction of every pixel of s's camera to dirs
// (width * height float4, w = 0): camera_ray, computed once instead of per path.
hipError_t launch_camera_dirs(const WfScene& s, float4* dirs, hipStream_t stream);
// First-bounce cache.  The primary ray of a pixel has no jitter, so its
// closest hit, the shading at that hit and the shadow ray from it towards
// each light are the same in every frame; only the light a frame picks and
// the direction it continues in are random.  One allocation of
// first_cache_bytes(pixels, lights) per camera and scene holds
//  * per pixel a 64-byte record (4 x uint4, floats as bits):
//      [0] closest sphere of the primary ray (-1: miss), its t, the lit mask
//          (bit l: light l reaches the hit; scenes of at most kFirstMaxLights
//          lights), the hit's material;
//      [1] normal x, y, z, origin of the rays leaving the hit x;
//      [2] origin y, z, radiance after the hit's emission r, g;
//      [3] radiance b, the pixel's index in the hit list (hits only), 0, 0;
//  * then, for scenes of 1..kFirstMaxLights lights, per light and pixel
//    ([light][pixel]) the light's Cook-Torrance term at the hit (float4,
//    w = 0; zero when the light does not reach the hit);
//  * then the list of pixels whose primary ray hits (uint32 each, in no
//    particular order) and its length (uint32, 16-byte slot).
// The path kernel then does bounce 0 from it where it hands out a job.
constexpr uint32_t kFirstMaxLights = 32;
struct FirstCacheView {
    uint4* rec;
    float4* light;
    uint32_t* hits;
    uint32_t* n_hits;
};
// Bytes of the hit list of a cache of `pixels` pixels (16-byte multiple);
// the list's length follows it.
__host__ __device__ inline size_t first_hits_bytes(size_t pixels)
{
    return (pixels * 4 + 15) & ~size_t(15);
}
// Jobs per frame of a table entry: the length of its hit list, or its pixels.
__host__ __device__ inline uint32_t job_slots(const uint32_t* hit_px, uint32_t pixels)
{
    return hit_px ? *(const uint32_t*)((const char*)hit_px + first_hits_bytes(pixels)) : pixels;
}
inline size_t first_light_bytes(size_t pixels, uint32_t lights)
{
    return lights >= 1 && lights <= kFirstMaxLights ? pixels * lights * 16 : 0;
}
inline size_t first_cache_bytes(size_t pixels, uint32_t lights)
{
    return pixels * 64 + first_light_bytes(pixels, lights) + first_hits_bytes(pixels) + 16;
}
inline FirstCacheView first_cache_view(void* buf, size_t pixels, uint32_t lights)
{
    char* b = (char*)buf;
    FirstCacheView v;
    v.rec = (uint4*)b;
    v.light = first_light_bytes(pixels, lights) ? (float4*)(b + pixels * 64) : nullptr;
    v.hits = (uint32_t*)(b + pixels * 64 + first_light_bytes(pixels, lights));
    v.n_hits = (uint32_t*)((char*)v.hits + first_hits_bytes(pixels));
    return v;
}
// Points s.first / s.first_light into buf (first_cache_bytes of s's pixels and lights).
inline void set_first_cache(WfScene& s, void* buf)
{
    const FirstCacheView v = first_cache_view(buf, (size_t)s.width * s.height, s.n_light);
    s.first = v.rec;
    s.first_light = v.light;
}
// Writes the cache s.first points to (s.dirs must hold the primary
// directions), the hit list included.
hipError_t launch_first_hits(const WfScene& s, hipStream_t stream);
// Max resident 256-lane workgroups of the path kernel per CU (the instance
// wf_waves(waves, flat_only)).
int wavefront_max_blocks_per_cu(bool flat_only, uint32_t waves = 0);

}  // namespace atx
