// render_wavefront.hip -- persistent, path-regenerating renderer.
//
// The per-pixel path (Renderer.cu:287-387) runs as a per-lane state machine
// so that all lanes of a wave trace a ray in every iteration:
//   START   begin the path of the lane's (frame, pixel) job
//   CONT    trace the primary/continuation ray; shade the hit; pick a light
//   SHADOW  trace the shadow ray; add the light's Cook-Torrance contribution
//   (then Russian roulette + next direction, or the path ends)
// Every (frame, pixel) path is an independent job.  A lane whose path ends
// stores the sample in the launch's sample buffer and takes the next job at
// once, so the work is shared out one path at a time: no lane waits for
// another pixel's earlier frame and no wave is left with a long tail of
// serial frames when the queue runs dry.  k_accumulate then adds every
// pixel's samples in frame order (Renderer.cu:164-168).  The floating-point
// operations of one path are exactly the plain kernel's (and the oracle's)
// and the additions happen in the same order, so the results are
// bit-identical to rendering each scene alone, one frame per launch, with
// the reference's per-pixel loop.
//
// Work distribution: a wave takes blocks of `chunk_px` consecutive jobs of
// one scene from the launch's global counter and hands them to its lanes as
// they fall idle.  A scene's jobs come in groups of 8 neighbouring pixels
// times the launch's frames, frame after frame inside a group
// (job_pixel_frame), so a wave traces 8 neighbouring pixels in many frames
// at once: their rays start from the same few hits, which keeps the lanes of
// a wave on nearby tree nodes and makes whole waves miss the same spheres in
// the flat loop.  Samples are stored in job order: a wave's block of jobs
// fills whole 128-byte lines of the sample buffer, which no other wave
// touches, and k_accumulate reads a pixel's frames 8 samples apart.
//
// First-bounce cache: a pixel's primary ray, the shading at its hit and the
// shadow ray towards each light are the same in every frame.  When the
// launch has the cache (k_first_hit, once per camera and scene), bounce 0
// of a path is done where the job is handed out -- the cached term of the
// frame's light if it reaches the hit, roulette, the next direction -- and
// the lane enters the trace loop at its first continuation ray or ends the
// path at once.  With the cache's hit list, jobs cover only the pixels
// whose primary ray hits; k_accumulate adds the others' constant sample.
//
// Spheres and the scene table
// are read with wave-uniform addresses through the constant address space,
// which the compiler lowers to scalar loads (two spheres per s_load_dwordx8,
// used directly as VALU operands).  A wave works on one scene at a time and
// drains its lanes before it switches.
//
// Registers: the kernel is compiled twice, for 6 and for 7 resident waves
// per SIMD (amdgpu_waves_per_eu: at most 80 and 72 VGPRs; the second
// spills a few to scratch).  More waves hide more latency than the spills cost: launches that
// traverse trees take the 6-wave instance, launches whose scenes all use the
// flat loop the 7-wave one (the natural allocation, 93 VGPRs, allows 5).
#include "device_types.h"
#include "trace_bvh.h"
#include "trace_device.h"
#include "wavefront.h"

#include <stdlib.h>

namespace atx {
namespace {

constexpr uint32_t kIdle = 0, kCont = 1, kShadow = 2;

// The scene table and the sphere arrays do not change during a launch; read
// through the constant address space, wave-uniform loads from them become
// scalar loads (s_load) into SGPRs even though the kernel stores to global
// memory elsewhere.  (The host pass only parses the kernel; it gets no
// address space.)
#if defined(__HIP_DEVICE_COMPILE__)
#define ATX_CONST __attribute__((address_space(4)))
#else
#define ATX_CONST
#endif
using CScene = const ATX_CONST WfScene;
using CF4 = const ATX_CONST float4;

template <class T>
__device__ __forceinline__ const ATX_CONST T* as_const(const T* p)
{
    return (const ATX_CONST T*)(uintptr_t)p;
}

struct SphSrc {
    CF4* p;
    __device__ __forceinline__ float4 operator()(uint32_t i) const { return p[i]; }
};

__device__ __forceinline__ f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// Adds the wave's per-lane test counts to the launch's counters and clears
// them (every lane of the wave calls it).
__device__ __forceinline__ void flush_counters(unsigned long long* ctr, uint32_t lane, uint32_t& tests,
                                               uint32_t& boxes)
{
    const unsigned long long ts = wave_sum(tests);
    const unsigned long long bx = wave_sum(boxes);
    if (lane == 0) {
        if (ts) atomicAdd(ctr + kCtrTests, ts);
        if (bx) atomicAdd(ctr + kCtrBoxes, bx);
    }
    tests = boxes = 0u;
}

// ---- one bounce of the per-pixel path (Renderer.cu:319-381), in pieces
// shared by the path kernel and the first-bounce cache, so that both
// evaluate the same expressions in the same order ----

// Hit point, normal and the origin of the rays that leave the hit
// (Renderer.cu:325-333).
__device__ __forceinline__ void hit_frame(float4 sphere, const Hit& h, f3 ro, f3 rd, f3& pos, f3& nrm, f3& so)
{
    const f3 center = xyz(sphere);
    pos = add(sub(ro, center), scale(rd, h.t));
    nrm = normalize(pos);
    pos = add(pos, center);
    so = add(pos, scale(nrm, 0.0001f));
}

// Direction (normalised) and squared distance from pos to a light
// (Renderer.cu:343-346).
__device__ __forceinline__ void light_ray(float4 light_pos, f3 pos, f3& L, float& dist2)
{
    L = sub(xyz(light_pos), pos);
    dist2 = dot(L, L);
    L = normalize(L);
}

// The shadow decision on the shadow ray's closest hit (Renderer.cu:348-350):
// lit unless that hit lies strictly between the surface and the light.
__device__ __forceinline__ bool lit_by(const Hit& h, float dist2)
{
    const float st = h.id < 0 ? -1.0f : h.t;
    return !(st > 0.0f && st * st < dist2);
}

// The hit material's emission, added to col (Renderer.cu:334-337).
__device__ __forceinline__ void add_emission(const float4* mat, int32_t mid, f3 tp, f3& col)
{
    const float4 m2 = mat[3 * mid + 2];
    if (m2.w > 0.0f)
        col = add(col, mul(scale(xyz(m2), m2.w), tp));
}

// Emission of a continuation hit and the light its shadow ray goes to.
// Returns true when a shadow ray follows (the scene has lights).
__device__ __forceinline__ bool shade_hit(CScene* S, const Hit& h, f3 ro, f3 rd, f3 tp, uint32_t seed, f3& col,
                                          f3& nrm, f3& so, int32_t& mid, uint32_t& lidx, f3& L, float& dist2)
{
    f3 pos;
    hit_frame(as_const(S->sph)[h.id], h, ro, rd, pos, nrm, so);
    mid = S->sph_mat[h.id];
    add_emission(S->mat, mid, tp, col);
    if (S->n_light == 0)
        return false;
    lidx = umod(pcg_hash(seed), S->div_light);   // pcg_hash(seed) % n_light
    light_ray(S->light[2 * lidx], pos, L, dist2);
    return true;
}

// Cook-Torrance term of a light that reaches the hit, weighted by the path
// throughput (Renderer.cu:351-360).
__device__ __forceinline__ f3 light_term(const float4* mat, const float4* light, int32_t mid, uint32_t lidx, f3 nrm,
                                         f3 rd, f3 L, f3 tp)
{
    const float4 m0 = mat[3 * mid + 0];
    const float4 m1 = mat[3 * mid + 1];
    const f3 albedo = xyz(m0);
    const float metallic = m1.w;
    const f3 base = add(scale(xyz(m1), 1.0f - metallic), scale(albedo, metallic));
    const f3 f = cook_torrance(albedo, base, metallic, m0.w, nrm, neg(rd), L);
    const float4 l0 = light[2 * lidx + 0];
    const float4 l1 = light[2 * lidx + 1];
    return mul(mul(scale(xyz(l1), l0.w), f), tp);
}

__device__ __forceinline__ void add_light(CScene* S, int32_t mid, uint32_t lidx, f3 nrm, f3 rd, f3 L, f3 tp,
                                          f3& col)
{
    col = add(col, light_term(S->mat, S->light, mid, lidx, nrm, rd, L, tp));
}

// Russian roulette and the next direction (Renderer.cu:362-381).  Returns
// true when the path ends here.
__device__ __forceinline__ bool next_bounce(CScene* S, int32_t mid, f3 nrm, f3 so, f3& tp, f3& ro, f3& rd,
                                            int32_t& bounce, uint32_t& seed)
{
    const float4 m0 = S->mat[3 * mid + 0];
    const float4 m1 = S->mat[3 * mid + 1];
    tp = mul(tp, xyz(m0));
    ro = so;
    const float p = gmax(0.1f, gmin(1.0f, length(tp)));
    if (pcg_float(seed) > p)
        return true;
    tp = divs(tp, p);
    rd = sample_next(nrm, m1.w > 0.0f, m0.w, seed);
    ++bounce;
    if (bounce >= S->max_bounces)
        return true;
    seed += (uint32_t)bounce;
    return false;
}

// The pixel of the scene's job `job` (`slots` jobs per frame): the path
// state does not carry it, only the debug captures need it.
__device__ __forceinline__ uint32_t job_pixel(CScene* S, uint32_t slots, uint32_t job)
{
    uint32_t slot, f;
    job_pixel_frame(job, slots, S->div_group, slot, f);
    return S->hit_px ? S->hit_px[slot] : slot;
}

// The sample of the scene's job `job` (`slots` jobs per frame), and the
// last frame's debug capture.
__device__ __forceinline__ void store_sample(CScene* S, uint32_t slots, uint32_t job, f3 col)
{
    S->samples[job] = make_float3(col.x, col.y, col.z);
    if (S->dbg_radiance) {
        uint32_t slot, f;
        job_pixel_frame(job, slots, S->div_group, slot, f);
        if (f == S->frames - 1u)
            S->dbg_radiance[S->hit_px ? S->hit_px[slot] : slot] = make_float4(col.x, col.y, col.z, 1.0f);
    }
}

template <uint32_t kWaves>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kWaves))) void k_render_wavefront(WfPass a)
{
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lanes_below = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    const uint32_t P = a.chunk_px;
    // The instances for 7 and 8 resident waves per SIMD park the path state
    // that a trace does not read (17 values per lane) in LDS while the trace
    // runs: the state then needs no VGPRs there, and the trace fits the
    // smaller register budget without spilling (17 KB of LDS per
    // 256-lane workgroup, 136 KB per CU at 8 waves per SIMD).
    constexpr bool kPark = kWaves >= 7;
    constexpr uint32_t kParked = 17;
    __shared__ uint32_t parked_all[4][kPark ? kParked : 1][64];
    uint32_t(*parked)[64] = parked_all[threadIdx.x >> 6];

    // ---- wave-uniform scheduling state ----
    CScene* scenes = as_const(a.scenes);
    int32_t sid = -1;
    CScene* S = nullptr;
    int32_t pend = -1;              // a block of another scene, waiting for the lanes to drain
    bool qdone = false;
    bool first_fetch = true;
    const uint32_t n_waves = gridDim.x * 4u;
    const uint32_t wave_id = (uint32_t)__builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4u + (threadIdx.x >> 6)));
    uint32_t blk0 = 0, blk_next = 0, blk_end = 0;   // jobs left: [blk0 + next, blk0 + end) of scene S
    uint32_t slots = 0;                             // S's jobs per frame (job_slots)

    // ---- per-lane path state ----
    uint32_t state = kIdle;
    uint32_t job = 0, seed = 0;   // the scene's job index, PCG state
    int32_t bounce = 0;
    // (ro, rd): the lane's next ray -- the primary or continuation ray in
    // kCont, the shadow ray (from the hit towards the light) in kShadow;
    // vd: in kShadow, the direction of the ray that made the hit (the view
    // direction of the light's Cook-Torrance term)
    f3 ro = mk(0.0f, 0.0f, 0.0f), rd = ro, vd = ro, tp = ro, col = ro;
    f3 nrm = ro;                    // normal at the current hit
    float dist2 = 0.0f;             // in kShadow: squared distance to the light
    int32_t mid = 0;                // material of the current hit
    uint32_t lidx = 0;              // light of the current shadow ray
    bool capture = false;           // this path is the last frame's: write the debug capture
    // statistics: rays and cached rays per wave; test counts per lane, added
    // to the launch's 64-bit counters at the end (and earlier if one could
    // overflow); exact retraces (rare) straight to their counter
    unsigned long long rays = 0, reused = 0;
    uint32_t tests = 0, boxes = 0;

    for (;;) {
        // ---------------- hand jobs to idle lanes (uniform control flow) ----------------
        for (;;) {
            const unsigned long long idle = __ballot(state == kIdle);
            if (idle == 0ull)
                break;
            // Jobs are handed out once refill_min lanes wait (or all do):
            // the hand-out and bounce 0 from the cache (about a sixth of the
            // kernel's time) then run for many lanes at once instead of for
            // the one or two whose paths just ended, while the waiting lanes
            // cost the trace loop little (its length is set by its slowest
            // lane).  1080p, A/B: 16 lanes 35.0 -> 34.0 ms per 256-frame
            // launch, 3.72 -> 3.44 ms per 20-frame one.
            if (idle != ~0ull && (uint32_t)__popcll(idle) < a.refill_min && blk_next < blk_end)
                break;
            if (blk_next < blk_end) {
                const uint32_t avail = blk_end - blk_next;
                const uint32_t rank = (uint32_t)__popcll(idle & lanes_below);
                const uint32_t nidle = (uint32_t)__popcll(idle);
                uint32_t cached_rays = 0;   // this lane's rays taken from the first-bounce cache
                if (state == kIdle && rank < avail) {
                    job = blk0 + blk_next + rank;
                    uint32_t px, frame;
                    job_pixel_frame(job, slots, S->div_group, px, frame);
                    if (S->hit_px)
                        px = S->hit_px[px];
                    frame += S->frame0;
                    ro = mk(S->cam_pos[0], S->cam_pos[1], S->cam_pos[2]);
                    capture = (S->dbg_hit || S->dbg_shadow) && frame == S->frame0 + S->frames - 1u;
                    // bounce 0 from the first-bounce cache needs no primary
                    // direction unless the shadow ray is traced (too many lights)
                    const bool cached = S->first && !capture && S->max_bounces > 0;
                    if (cached && S->n_light <= kFirstMaxLights) {
                        rd = mk(0.0f, 0.0f, 0.0f);
                    } else if (S->dirs) {
                        const float4 dv = S->dirs[px];
                        rd = mk(dv.x, dv.y, dv.z);
                    } else {
                        const uint32_t W = S->width;
                        const uint32_t y = px / W, x = px - y * W;
                        float ip[16], iv[16];
                        for (int k = 0; k < 16; ++k) {
                            ip[k] = S->inv_proj[k];
                            iv[k] = S->inv_view[k];
                        }
                        rd = camera_ray(ip, iv, x, y, W, S->height);
                    }
                    tp = mk(1.0f, 1.0f, 1.0f);
                    col = mk(0.0f, 0.0f, 0.0f);
                    seed = px;   // x + y * W
                    seed *= frame;
                    bounce = 0;
                    if (capture) {
                        for (uint32_t i = 0; i < S->dbg_bounces; ++i) {
                            if (S->dbg_hit) S->dbg_hit[(size_t)px * S->dbg_bounces + i] = -2;
                            if (S->dbg_shadow) S->dbg_shadow[(size_t)px * S->dbg_bounces + i] = -2;
                        }
                    }
                    state = kCont;
                    if (cached) {
                        // bounce 0 from the first-bounce cache; the lane then enters
                        // the trace loop at its first continuation ray, or ends the
                        // path and takes another job
                        // the frame's light is known before the record arrives: issue
                        // the record's four loads and the light term's load together
                        const uint32_t nl = S->n_light;
                        const uint32_t lsel = nl > 0u ? umod(pcg_hash(seed), S->div_light) : 0u;
                        const uint4* fr = S->first + 4 * (size_t)px;
                        const uint4 r0 = fr[0], r1 = fr[1], r2 = fr[2], r3 = fr[3];
                        const float4 lt = S->first_light ? S->first_light[(size_t)lsel * (S->width * S->height) + px]
                                                         : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        const int32_t hid = (int32_t)r0.x;
                        cached_rays = 1;
                        bool done = false;
                        if (hid < 0) {
                            if (S->sky_light)
                                col = add(col, mul(mk(0.6f, 0.7f, 0.9f), tp));
                            done = true;
                        } else if (S->n_light > kFirstMaxLights) {
                            // too many lights for the mask: shade the cached hit and
                            // trace the shadow ray
                            Hit h;
                            h.id = hid;
                            h.t = __uint_as_float(r0.y);
                            f3 so, L;
                            shade_hit(S, h, ro, rd, tp, seed, col, nrm, so, mid, lidx, L, dist2);
                            vd = rd;
                            ro = so;
                            rd = L;
                            state = kShadow;
                        } else {
                            nrm = mk(__uint_as_float(r1.x), __uint_as_float(r1.y), __uint_as_float(r1.z));
                            const f3 so = mk(__uint_as_float(r1.w), __uint_as_float(r2.x), __uint_as_float(r2.y));
                            col = mk(__uint_as_float(r2.z), __uint_as_float(r2.w), __uint_as_float(r3.x));
                            mid = (int32_t)r0.w;
                            if (nl > 0u) {
                                // the frame's light, and its term if it reaches the hit
                                lidx = lsel;
                                cached_rays = 2;
                                if ((r0.z >> lidx) & 1u)
                                    col = add(col, xyz(lt));
                            }
                            done = next_bounce(S, mid, nrm, so, tp, ro, rd, bounce, seed);
                        }
                        if (done) {
                            store_sample(S, slots, job, col);
                            state = kIdle;
                        }
                    }
                }
                reused += (unsigned long long)(__popcll(__ballot(cached_rays >= 1u)) +
                                               __popcll(__ballot(cached_rays >= 2u)));
                blk_next += nidle < avail ? nidle : avail;
                continue;
            }
            // the block is used up: take the next one
            int32_t c = pend;
            if (c < 0) {
                if (qdone)
                    break;
                // a wave's first chunk is its own index in the grid; the
                // counter hands out the chunks after the grid's first
                // n_waves (all waves asking the one counter at once made the
                // first fetch wait 35 us on average)
                uint32_t got = 0;
                if (first_fetch) {
                    got = wave_id;
                    first_fetch = false;
                } else {
                    if (lane == 0)
                        got = atomicAdd(a.queue, 1u) + n_waves;
                    got = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(got, 0, 64));
                }
                if (got >= a.total_chunks) {
                    qdone = true;
                    break;
                }
                c = (int32_t)got;
            }
            int32_t s = sid < 0 ? 0 : sid;
            while (s + 1 < (int32_t)a.n_scenes && (uint32_t)c >= scenes[s].chunk_begin + scenes[s].n_chunks)
                ++s;
            if (s != sid) {
                // a new scene: only once every lane has finished its path
                if (idle != ~0ull) {
                    pend = c;
                    break;
                }
                sid = s;
                S = &scenes[s];
                slots = (uint32_t)__builtin_amdgcn_readfirstlane((int)job_slots(S->hit_px, S->width * S->height));
            }
            pend = -1;
            const uint32_t jobs = slots * S->frames;
            blk0 = ((uint32_t)c - S->chunk_begin) * P;
            blk_end = blk0 >= jobs ? 0u : (jobs - blk0 < P ? jobs - blk0 : P);
            blk_next = 0;
            // the host may lay a scene out for more jobs than its hit list
            // holds: past the last scene's jobs every later chunk is empty
            if (blk_end == 0u && s + 1 == (int32_t)a.n_scenes)
                qdone = true;
        }
        if (!__any(state != kIdle) && qdone && pend < 0)
            break;

        // ---------------- one closest-hit trace for every tracing lane ----------------
        const bool tracing = state == kCont || state == kShadow;
        bool end_bounce = false;   // run RR + next direction
        bool end_path = state == kCont && S->max_bounces <= 0;   // zero bounces: black sample
        const bool traces = tracing && !end_path;
        rays += (unsigned long long)__popcll(__ballot(traces));
        if (traces) {
            const bool shadow = state == kShadow;
            const SphSrc flat{as_const(S->sph)};
            Hit h;
            if constexpr (kPark) {
                const uint32_t v[kParked] = {__float_as_uint(col.x), __float_as_uint(col.y), __float_as_uint(col.z),
                                             __float_as_uint(tp.x),  __float_as_uint(tp.y),  __float_as_uint(tp.z),
                                             __float_as_uint(nrm.x), __float_as_uint(nrm.y), __float_as_uint(nrm.z),
                                             __float_as_uint(vd.x),  __float_as_uint(vd.y),  __float_as_uint(vd.z),
                                             seed, (uint32_t)bounce, job, (uint32_t)mid, lidx};
                for (uint32_t k = 0; k < kParked; ++k)
                    parked[k][lane] = v[k];
                // keep the compiler from forwarding the stored values past the trace
                __asm__ volatile("" ::: "memory");
            }
            uint32_t dt = 0, db = 0;   // the trace's sphere and box tests
            if (S->bvh.n_nodes) {
                // the debug capture records the closest occluder; otherwise any proven one will do
                const BvhDev B = S->bvh;
                BvhQuery q;
                const float qd2 = (shadow && !(capture && S->dbg_shadow)) ? dist2 : -1.0f;
                h = trace_bvh<true>(B, flat, S->n_sph, ro, rd, qd2, q);
                if (q.exact)
                    atomicAdd(a.counters + kCtrExact, (unsigned long long)q.exact);
                dt = q.tests;
                db = q.boxes;
            } else {
                uint32_t ex = 0;
                h = trace_filtered(flat, S->n_sph, ro, rd, ex);
                if (ex)
                    atomicAdd(a.counters + kCtrExact, (unsigned long long)ex);
                dt = S->n_sph;
            }
            if constexpr (kPark) {
                __asm__ volatile("" ::: "memory");
                col = mk(__uint_as_float(parked[0][lane]), __uint_as_float(parked[1][lane]),
                         __uint_as_float(parked[2][lane]));
                tp = mk(__uint_as_float(parked[3][lane]), __uint_as_float(parked[4][lane]),
                        __uint_as_float(parked[5][lane]));
                nrm = mk(__uint_as_float(parked[6][lane]), __uint_as_float(parked[7][lane]),
                         __uint_as_float(parked[8][lane]));
                if (shadow)
                    vd = mk(__uint_as_float(parked[9][lane]), __uint_as_float(parked[10][lane]),
                            __uint_as_float(parked[11][lane]));
                seed = parked[12][lane];
                bounce = (int32_t)parked[13][lane];
                job = parked[14][lane];
                mid = (int32_t)parked[15][lane];
                lidx = parked[16][lane];
            }
            tests += dt;
            boxes += db;

            // ---------------- shading (divergent) ----------------
            if (!shadow) {
                if (capture && S->dbg_hit && (uint32_t)bounce < S->dbg_bounces)
                    S->dbg_hit[(size_t)job_pixel(S, slots, job) * S->dbg_bounces + bounce] = h.id;
                if (h.id < 0) {
                    if (S->sky_light)
                        col = add(col, mul(mk(0.6f, 0.7f, 0.9f), tp));
                    end_path = true;
                } else {
                    f3 so, L;
                    if (shade_hit(S, h, ro, rd, tp, seed, col, nrm, so, mid, lidx, L, dist2)) {
                        vd = rd;
                        rd = L;
                        state = kShadow;
                    } else {
                        end_bounce = true;
                    }
                    ro = so;
                }
            } else {
                if (capture && S->dbg_shadow && (uint32_t)bounce < S->dbg_bounces)
                    S->dbg_shadow[(size_t)job_pixel(S, slots, job) * S->dbg_bounces + bounce] = h.id;
                if (lit_by(h, dist2))
                    add_light(S, mid, lidx, nrm, vd, rd, tp, col);
                end_bounce = true;
            }
        }
        if (end_bounce) {
            // the next ray leaves from the hit's offset origin, which ro holds
            if (next_bounce(S, mid, nrm, ro, tp, ro, rd, bounce, seed))
                end_path = true;
            else
                state = kCont;
        }
        if (__any(tests >= 0x80000000u || boxes >= 0x80000000u))
            flush_counters(a.counters, lane, tests, boxes);
        if (end_path) {
            store_sample(S, slots, job, col);
            state = kIdle;
        }
    }

    flush_counters(a.counters, lane, tests, boxes);
    if (lane == 0 && rays)
        atomicAdd(a.counters + kCtrRays, rays);
    if (lane == 0 && reused)
        atomicAdd(a.counters + kCtrReused, reused);
}

// Adds every pixel's samples of the launch to its accumulation in frame order
// (Renderer.cu:164-165; frame 1 starts from zero, which is what the
// reference's memset provides) and writes the RGBA8 value (Renderer.cu:166-168,
// Renderer.h:70-78).  One lane per pixel of the launch's pixel range; a
// pixel's samples lie at job_of(slot, f), m samples apart from one frame to
// the next (m = 8 but in a scene's last group of slots).
__global__ __launch_bounds__(256) void k_accumulate(WfPass a)
{
    const uint32_t g = blockIdx.x * 256u + threadIdx.x;
    if (g >= a.total_px)
        return;
    uint32_t lo = 0, hi = a.n_scenes;          // last scene with px_begin <= g
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) / 2u;
        if (a.scenes[mid].px_begin <= g) lo = mid;
        else hi = mid;
    }
    const WfScene& S = a.scenes[lo];
    const uint32_t npx = S.width * S.height;
    const uint32_t px = g - S.px_begin;
    if (px >= npx)
        return;
    const bool first = S.frame0 == 1u;
    // a pixel left out of the jobs (its primary ray misses): every frame's
    // sample is the one such a path returns, the sky colour or black
    const bool sky = S.hit_px && (int32_t)S.first[4 * (size_t)px].x < 0;
    const float4 sky_c = S.sky_light ? make_float4(0.6f, 0.7f, 0.9f, 1.0f) : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    float4 acc = first ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : S.accum[px];
    // this pixel's samples: frame f at sp[f * step] (slots are hit-list
    // indices with a hit list, else pixels; a sky pixel has none)
    const uint32_t slots = job_slots(S.hit_px, npx);
    const uint32_t slot = sky ? 0u : S.hit_px ? S.first[4 * (size_t)px + 3].y : px;
    const uint32_t left = slots - (slot & ~(kJobGroupPx - 1u));
    const uint32_t step = left < kJobGroupPx ? left : kJobGroupPx;
    const ATX_GLOBAL float3* sp = (const ATX_GLOBAL float3*)(uintptr_t)S.samples + job_of(slot, 0u, slots, S.frames);
    if (S.kahan) {
        float4 cp = first ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : S.comp[px];
        const auto add = [&](const float4 c) {
            float yv, t;
            yv = c.x - cp.x; t = acc.x + yv; cp.x = (t - acc.x) - yv; acc.x = t;
            yv = c.y - cp.y; t = acc.y + yv; cp.y = (t - acc.y) - yv; acc.y = t;
            yv = c.z - cp.z; t = acc.z + yv; cp.z = (t - acc.z) - yv; acc.z = t;
            yv = c.w - cp.w; t = acc.w + yv; cp.w = (t - acc.w) - yv; acc.w = t;
        };
        // two loops: a select between sky_c and a loaded sample becomes a
        // select of addresses, and sky_c then lives in scratch memory
        if (sky) {
            for (uint32_t f = 0; f < S.frames; ++f)
                add(sky_c);
        } else {
            for (uint32_t f = 0; f < S.frames; ++f) {
                const float3 c = sp[(size_t)f * step];
                add(make_float4(c.x, c.y, c.z, 1.0f));
            }
        }
        S.comp[px] = cp;
    } else if (sky) {
        for (uint32_t f = 0; f < S.frames; ++f) {
            acc.x = acc.x + sky_c.x;
            acc.y = acc.y + sky_c.y;
            acc.z = acc.z + sky_c.z;
            acc.w = acc.w + sky_c.w;
        }
    } else {
        // four frames' loads in flight ahead of their additions (in frame
        // order); every sample's alpha is 1
        uint32_t f = 0;
        for (; f + 4u <= S.frames; f += 4u) {
            float3 c[4];
            for (int k = 0; k < 4; ++k)
                c[k] = sp[(size_t)(f + k) * step];
            for (int k = 0; k < 4; ++k) {
                acc.x = acc.x + c[k].x;
                acc.y = acc.y + c[k].y;
                acc.z = acc.z + c[k].z;
                acc.w = acc.w + 1.0f;
            }
        }
        for (; f < S.frames; ++f) {
            const float3 c = sp[(size_t)f * step];
            acc.x = acc.x + c.x;
            acc.y = acc.y + c.y;
            acc.z = acc.z + c.z;
            acc.w = acc.w + 1.0f;
        }
    }
    S.accum[px] = acc;
    const float fi = (float)(S.frame0 + S.frames - 1u);
    S.image[px] = rgba8(gclamp(acc.x / fi, 0.0f, 1.0f), gclamp(acc.y / fi, 0.0f, 1.0f),
                        gclamp(acc.z / fi, 0.0f, 1.0f), gclamp(acc.w / fi, 0.0f, 1.0f));
}

struct CamDirArgs {
    float inv_proj[16];
    float inv_view[16];
    uint32_t width, height;
    float4* dirs;
};

__global__ __launch_bounds__(256) void k_camera_dirs(CamDirArgs c)
{
    const uint32_t px = blockIdx.x * 256u + threadIdx.x;
    if (px >= c.width * c.height)
        return;
    const uint32_t y = px / c.width, x = px - y * c.width;
    const f3 d = camera_ray(c.inv_proj, c.inv_view, x, y, c.width, c.height);
    c.dirs[px] = make_float4(d.x, d.y, d.z, 0.0f);
}

struct FirstArgs {
    WfScene s;
    uint4* rec;        // 4 per pixel
    float4* light;     // [light][pixel], or nullptr
    uint32_t* hits;    // pixels whose primary ray hits, appended
    uint32_t* n_hits;  // their count (zero before the launch)
};

// One lane per pixel: the primary ray's closest hit, the shading at it, and
// for scenes of at most kFirstMaxLights lights each light's shadow decision
// and Cook-Torrance term, with the path kernel's expressions (hit_frame,
// add_emission, light_ray, lit_by, light_term; tp = 1 at bounce 0) and
// queries, so the cached values are the ones every frame's path computes.
// A wave takes an 8 x 8 tile of pixels: its rays stay closer together than
// along a 64-pixel row (1080p: 0.586 -> 0.556 ms), and the hit list then
// runs tile by tile.
__global__ __launch_bounds__(256) void k_first_hit(FirstArgs a)
{
    const WfScene& s = a.s;
    const uint32_t npx = s.width * s.height;
    // one 8 x 8 tile of pixels per wave
    const uint32_t tiles_x = (s.width + 7u) / 8u;
    const uint32_t wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const uint32_t ty = wave / tiles_x, tx = wave - ty * tiles_x;
    const uint32_t x = tx * 8u + (threadIdx.x & 7u), y = ty * 8u + ((threadIdx.x >> 3) & 7u);
    if (x >= s.width || y >= s.height)
        return;
    const uint32_t px = y * s.width + x;
    const f3 ro = mk(s.cam_pos[0], s.cam_pos[1], s.cam_pos[2]);
    const float4 dv = s.dirs[px];
    const f3 rd = mk(dv.x, dv.y, dv.z);
    const SphSrc flat{as_const(s.sph)};
    auto trace = [&](f3 o, f3 d, float d2) -> Hit {
        if (s.bvh.n_nodes) {
            BvhQuery q;
            return trace_bvh(s.bvh, flat, s.n_sph, o, d, d2, q);
        }
        uint32_t ex = 0;
        return trace_filtered(flat, s.n_sph, o, d, ex);
    };
    const Hit h = trace(ro, rd, -1.0f);
    uint32_t slot = 0u;   // index in the hit list
    {
        // append this wave's hit pixels to the list (one atomic per wave)
        const uint32_t lane = threadIdx.x & 63u;
        const unsigned long long act = __ballot(1);
        const unsigned long long hm = __ballot(h.id >= 0);
        const uint32_t leader = (uint32_t)__ffsll((long long)act) - 1u;
        uint32_t base = 0;
        if (lane == leader && hm)
            base = atomicAdd(a.n_hits, (uint32_t)__popcll(hm));
        base = (uint32_t)__shfl((int)base, (int)leader, 64);
        if (h.id >= 0) {
            const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64u - lane));
            slot = base + (uint32_t)__popcll(hm & below);
            a.hits[slot] = px;
        }
    }
    const f3 one = mk(1.0f, 1.0f, 1.0f);
    const f3 zero = mk(0.0f, 0.0f, 0.0f);
    f3 nrm = zero, so = zero, col = zero;
    int32_t mid = 0;
    uint32_t mask = 0u;
    if (h.id >= 0) {
        f3 pos;
        hit_frame(as_const(s.sph)[h.id], h, ro, rd, pos, nrm, so);
        mid = s.sph_mat[h.id];
        add_emission(s.mat, mid, one, col);
        if (a.light) {
            for (uint32_t l = 0; l < s.n_light; ++l) {
                f3 L;
                float dist2;
                light_ray(s.light[2 * l], pos, L, dist2);
                f3 term = zero;
                if (lit_by(trace(so, L, dist2), dist2)) {
                    mask |= 1u << l;
                    term = light_term(s.mat, s.light, mid, l, nrm, rd, L, one);
                }
                a.light[(size_t)l * npx + px] = make_float4(term.x, term.y, term.z, 0.0f);
            }
        }
    } else if (a.light) {
        for (uint32_t l = 0; l < s.n_light; ++l)
            a.light[(size_t)l * npx + px] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    uint4* r = a.rec + 4 * (size_t)px;
    r[0] = make_uint4((uint32_t)h.id, __float_as_uint(h.t), mask, (uint32_t)mid);
    r[1] = make_uint4(__float_as_uint(nrm.x), __float_as_uint(nrm.y), __float_as_uint(nrm.z), __float_as_uint(so.x));
    r[2] = make_uint4(__float_as_uint(so.y), __float_as_uint(so.z), __float_as_uint(col.x), __float_as_uint(col.y));
    r[3] = make_uint4(__float_as_uint(col.z), slot, 0u, 0u);
}

__global__ void k_zero_u32(unsigned int* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

hipError_t launch_wavefront(const WfPass& pass, const WfLaunch& l, hipStream_t stream, hipEvent_t paths_done)
{
    if (pass.total_px == 0)
        return hipSuccess;
    WfPass p = pass;
    static const uint32_t refill = [] {
        const char* env = getenv("ATX_WF_REFILL");
        const long v = env ? strtol(env, nullptr, 10) : (long)kRefillMin;
        return (uint32_t)(v < 1 ? 1 : v > 64 ? 64 : v);
    }();
    p.refill_min = refill;
    if (p.chunk_px == 0 || p.chunk_px % 64 != 0 || p.chunk_px > kMaxChunkPx)
        return hipErrorInvalidValue;
    // no jobs (every pixel's primary ray misses): k_accumulate alone
    if (p.total_chunks > 0) {
        switch (wf_waves(l.waves, l.flat_only)) {
        case 5: hipLaunchKernelGGL(k_render_wavefront<5>, dim3(l.workgroups), dim3(256), 0, stream, p); break;
        case 6: hipLaunchKernelGGL(k_render_wavefront<6>, dim3(l.workgroups), dim3(256), 0, stream, p); break;
        case 7: hipLaunchKernelGGL(k_render_wavefront<7>, dim3(l.workgroups), dim3(256), 0, stream, p); break;
        default: hipLaunchKernelGGL(k_render_wavefront<8>, dim3(l.workgroups), dim3(256), 0, stream, p); break;
        }
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && paths_done)
        e = hipEventRecord(paths_done, stream);
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_accumulate, dim3((p.total_px + 255u) / 256u), dim3(256), 0, stream, p);
    return hipGetLastError();
}

hipError_t launch_camera_dirs(const WfScene& s, float4* dirs, hipStream_t stream)
{
    const uint32_t n = s.width * s.height;
    if (n == 0)
        return hipSuccess;
    CamDirArgs c;
    for (int k = 0; k < 16; ++k) {
        c.inv_proj[k] = s.inv_proj[k];
        c.inv_view[k] = s.inv_view[k];
    }
    c.width = s.width;
    c.height = s.height;
    c.dirs = dirs;
    hipLaunchKernelGGL(k_camera_dirs, dim3((n + 255u) / 256u), dim3(256), 0, stream, c);
    return hipGetLastError();
}

hipError_t launch_first_hits(const WfScene& s, hipStream_t stream)
{
    const uint32_t n = s.width * s.height;
    if (n == 0)
        return hipSuccess;
    const bool lights = s.n_light >= 1 && s.n_light <= kFirstMaxLights;
    if (!s.dirs || !s.first || lights != (s.first_light != nullptr))
        return hipErrorInvalidValue;
    const FirstCacheView v = first_cache_view(const_cast<uint4*>(s.first), n, s.n_light);
    FirstArgs a;
    a.s = s;
    a.rec = v.rec;
    a.light = v.light;
    a.hits = v.hits;
    a.n_hits = v.n_hits;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, stream, v.n_hits, 1u);
    const uint32_t tiles = ((s.width + 7u) / 8u) * ((s.height + 7u) / 8u);
    hipLaunchKernelGGL(k_first_hit, dim3((tiles + 3u) / 4u), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_zero_u32(unsigned int* p, uint32_t n, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

int wavefront_max_blocks_per_cu(bool flat_only, uint32_t waves)
{
    int blocks = 0;
    hipError_t e;
    switch (wf_waves(waves, flat_only)) {
    case 5: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_render_wavefront<5>, 256, 0); break;
    case 6: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_render_wavefront<6>, 256, 0); break;
    case 7: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_render_wavefront<7>, 256, 0); break;
    default: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_render_wavefront<8>, 256, 0); break;
    }
    if (e != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

}  // namespace atx
