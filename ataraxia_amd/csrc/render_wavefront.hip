// render_wavefront.hip -- persistent, path-regenerating renderer.
//
// The per-pixel path (Renderer.cu:287-387) runs as a per-lane state machine
// so that all lanes of a wave trace a ray in every iteration:
//   WAIT    the lane holds sample (pixel, frame) but the pixel's previous
//           frame is still in flight in another lane
//   CONT    trace the primary/continuation ray; shade the hit; pick a light
//   SHADOW  trace the shadow ray; add the light's Cook-Torrance contribution
//   (then Russian roulette + next direction, or the path ends)
// A lane whose path ends accumulates its sample into the chunk's LDS
// accumulator and immediately takes the next sample of the wave's chunk.  The
// floating-point operations of one path are exactly the plain kernel's (and
// the oracle's) and each pixel's samples are added in frame order, so the
// results are bit-identical to rendering each scene alone, one frame per
// launch, with the reference's per-pixel loop.
//
// Work unit: a chunk = `chunk_px` consecutive pixels x all frames of the
// launch.  Each wave keeps two chunk slots in LDS (accumulation, optional
// Kahan compensation, last finished frame per pixel): while the lanes finish
// one chunk's last paths the next chunk's samples are already being handed
// out, so a wave only drains when it changes scene or the work runs out.
//
// Spheres are read with wave-uniform indices from global memory, which the
// compiler lowers to scalar loads (two spheres per s_load_dwordx8, used
// directly as VALU operands): no LDS staging, no barrier, waves independent.
#include "device_types.h"
#include "trace_bvh.h"
#include "trace_device.h"
#include "wavefront.h"

namespace atx {
namespace {

constexpr uint32_t kIdle = 0, kCont = 1, kShadow = 2, kWait = 3;

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

// LDS slot of one chunk: acc[P] float4, comp[P] float4 (Kahan only), done[P] u32.
struct SlotView {
    float4* acc;
    float4* comp;
    uint32_t* done;
};

__device__ __forceinline__ SlotView slot_view(char* wave_base, uint32_t k, uint32_t P, uint32_t kahan)
{
    const uint32_t slot_bytes = P * (20u + 16u * kahan);
    char* b = wave_base + k * slot_bytes;
    SlotView v;
    v.acc = reinterpret_cast<float4*>(b);
    v.comp = reinterpret_cast<float4*>(b + 16u * P);
    v.done = reinterpret_cast<uint32_t*>(b + 16u * P * (1u + kahan));
    return v;
}

__global__ __launch_bounds__(256) void k_render_wavefront(WfPass a)
{
    extern __shared__ __align__(16) char lds[];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t wave = threadIdx.x >> 6;
    const unsigned long long lanes_below = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
    const uint32_t P = a.chunk_px;
    const uint32_t K = a.kahan;
    char* wave_lds = lds + (size_t)wave * 2u * P * (20u + 16u * K);

    // ---- wave-uniform scheduling state ----
    CScene* scenes = as_const(a.scenes);
    int32_t sid = -1;
    CScene* S = nullptr;
    int32_t pend = -1;
    bool qdone = false;
    uint32_t cs = 0;                                   // slot whose samples are handed out
    bool sl_on[2] = {false, false};
    uint32_t sl_p0[2] = {0, 0}, sl_nv[2] = {0, 0};     // first pixel, valid pixels
    uint32_t sl_next[2] = {0, 0}, sl_total[2] = {0, 0};

    // ---- per-lane path state ----
    uint32_t state = kIdle, lslot = 0, li = 0, frame = 0;
    uint32_t px = 0, seed = 0;
    int32_t bounce = 0;
    f3 ro = mk(0.0f, 0.0f, 0.0f), rd = ro, tp = ro, col = ro;
    f3 so = ro, nrm = ro, L = ro;   // shadow/next origin, normal, light direction
    float dist2 = 0.0f;
    int32_t mid = 0;                // material of the current hit
    uint32_t lidx = 0;              // light of the current shadow ray
    uint32_t rays = 0, exact = 0;
    unsigned long long tests = 0, boxes = 0;

    for (;;) {
        // ---------------- write back chunks whose samples are all done ----------------
        for (uint32_t k = 0; k < 2; ++k) {
            if (!sl_on[k] || sl_next[k] != sl_total[k])
                continue;
            if (__any(state != kIdle && lslot == k))
                continue;
            const SlotView v = slot_view(wave_lds, k, P, K);
            const uint32_t flast = S->frame0 + S->frames - 1u;
            const float fi = (float)flast;
            for (uint32_t i = lane; i < sl_nv[k]; i += 64u) {
                const uint32_t q = sl_p0[k] + i;
                const float4 acc = v.acc[i];
                S->accum[q] = acc;
                if (S->kahan) S->comp[q] = v.comp[i];
                S->image[q] = rgba8(gclamp(acc.x / fi, 0.0f, 1.0f), gclamp(acc.y / fi, 0.0f, 1.0f),
                                    gclamp(acc.z / fi, 0.0f, 1.0f), gclamp(acc.w / fi, 0.0f, 1.0f));
            }
            sl_on[k] = false;
            __builtin_amdgcn_wave_barrier();
        }

        // ---------------- refill idle lanes (uniform control flow) ----------------
        for (;;) {
            const unsigned long long idle = __ballot(state == kIdle);
            if (idle == 0ull)
                break;
            if (sl_on[cs] && sl_next[cs] < sl_total[cs]) {
                const uint32_t avail = sl_total[cs] - sl_next[cs];
                const uint32_t rank = (uint32_t)__popcll(idle & lanes_below);
                const uint32_t nidle = (uint32_t)__popcll(idle);
                if (state == kIdle && rank < avail) {
                    const uint32_t t = sl_next[cs] + rank;
                    const uint32_t fo = t / sl_nv[cs];
                    li = t - fo * sl_nv[cs];
                    frame = S->frame0 + fo;
                    lslot = cs;
                    state = kWait;
                }
                sl_next[cs] += nidle < avail ? nidle : avail;
                continue;
            }
            // the current slot is exhausted: load the next chunk into a free slot
            uint32_t fs;
            if (!sl_on[cs]) fs = cs;
            else if (!sl_on[cs ^ 1u]) fs = cs ^ 1u;
            else break;   // both slots in flight: wait for a write-back
            int32_t c = pend;
            if (c < 0) {
                if (qdone)
                    break;
                uint32_t got = 0;
                if (lane == 0)
                    got = atomicAdd(a.queue, 1u);
                got = (uint32_t)__builtin_amdgcn_readfirstlane((int)__shfl(got, 0, 64));
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
                // a new scene: only once every lane and both slots are drained
                if (idle != ~0ull || sl_on[0] || sl_on[1]) {
                    pend = c;
                    break;
                }
                sid = s;
                S = &scenes[s];
            }
            pend = -1;
            const uint32_t npx = S->width * S->height;
            const uint32_t p0 = ((uint32_t)c - S->chunk_begin) * P;
            const uint32_t nv = npx - p0 < P ? npx - p0 : P;
            const SlotView v = slot_view(wave_lds, fs, P, K);
            const bool first = S->frame0 == 1u;   // frame 1 starts from zero (the reference's memset)
            for (uint32_t i = lane; i < nv; i += 64u) {
                v.acc[i] = first ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : S->accum[p0 + i];
                if (K) v.comp[i] = (first || !S->kahan) ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : S->comp[p0 + i];
                v.done[i] = S->frame0 - 1u;
            }
            __builtin_amdgcn_wave_barrier();
            sl_on[fs] = true;
            sl_p0[fs] = p0;
            sl_nv[fs] = nv;
            sl_next[fs] = 0;
            sl_total[fs] = nv * S->frames;
            cs = fs;
        }

        // ---------------- start samples whose pixel's previous frame is done ----------------
        if (state == kWait) {
            const SlotView v = slot_view(wave_lds, lslot, P, K);
            if (v.done[li] + 1u == frame) {
                px = sl_p0[0] + li;
                px = lslot ? sl_p0[1] + li : px;
                const uint32_t W = S->width;
                const uint32_t y = px / W, x = px - y * W;
                ro = mk(S->cam_pos[0], S->cam_pos[1], S->cam_pos[2]);
                float ip[16], iv[16];
                for (int k = 0; k < 16; ++k) {
                    ip[k] = S->inv_proj[k];
                    iv[k] = S->inv_view[k];
                }
                rd = camera_ray(ip, iv, x, y, W, S->height);
                tp = mk(1.0f, 1.0f, 1.0f);
                col = mk(0.0f, 0.0f, 0.0f);
                seed = x + y * W;
                seed *= frame;
                bounce = 0;
                if (S->dbg_hit || S->dbg_shadow) {
                    for (uint32_t i = 0; i < S->dbg_bounces; ++i) {
                        if (S->dbg_hit) S->dbg_hit[(size_t)px * S->dbg_bounces + i] = -2;
                        if (S->dbg_shadow) S->dbg_shadow[(size_t)px * S->dbg_bounces + i] = -2;
                    }
                }
                state = kCont;
            }
        }
        if (!__any(state != kIdle) && !sl_on[0] && !sl_on[1] && qdone && pend < 0)
            break;

        // ---------------- one closest-hit trace for every tracing lane ----------------
        const bool tracing = state == kCont || state == kShadow;
        bool end_bounce = false;   // run RR + next direction
        bool end_path = state == kCont && S->max_bounces <= 0;   // zero bounces: black sample
        if (tracing && !end_path) {
            const bool shadow = state == kShadow;
            const f3 o = shadow ? so : ro;
            const f3 d = shadow ? L : rd;
            const SphSrc flat{as_const(S->sph)};
            Hit h;
            if (S->bvh.n_nodes) {
                // the debug capture records the closest occluder; otherwise any proven one will do
                const BvhDev B = S->bvh;
                BvhQuery q;
                const float qd2 = (shadow && S->dbg_shadow == nullptr) ? dist2 : -1.0f;
                h = trace_bvh(B, flat, S->n_sph, o, d, qd2, q);
                exact += q.exact;
                tests += q.tests;
                boxes += q.boxes;
            } else {
                h = trace_filtered(flat, S->n_sph, o, d, exact);
                tests += S->n_sph;
            }
            ++rays;

            // ---------------- shading (divergent) ----------------
            if (!shadow) {
                if (S->dbg_hit && (uint32_t)bounce < S->dbg_bounces)
                    S->dbg_hit[(size_t)px * S->dbg_bounces + bounce] = h.id;
                if (h.id < 0) {
                    if (S->sky_light)
                        col = add(col, mul(mk(0.6f, 0.7f, 0.9f), tp));
                    end_path = true;
                } else {
                    const f3 center = xyz(as_const(S->sph)[h.id]);
                    f3 pos = add(sub(ro, center), scale(rd, h.t));
                    nrm = normalize(pos);
                    pos = add(pos, center);
                    mid = S->sph_mat[h.id];
                    const float4 m2 = S->mat[3 * mid + 2];
                    if (m2.w > 0.0f)
                        col = add(col, mul(scale(xyz(m2), m2.w), tp));
                    so = add(pos, scale(nrm, 0.0001f));
                    if (S->n_light > 0) {
                        lidx = pcg_hash(seed) % S->n_light;
                        L = sub(xyz(S->light[2 * lidx]), pos);
                        dist2 = dot(L, L);
                        L = normalize(L);
                        state = kShadow;
                    } else {
                        end_bounce = true;
                    }
                }
            } else {
                if (S->dbg_shadow && (uint32_t)bounce < S->dbg_bounces)
                    S->dbg_shadow[(size_t)px * S->dbg_bounces + bounce] = h.id;
                const float st = h.id < 0 ? -1.0f : h.t;
                if (!(st > 0.0f && st * st < dist2)) {
                    const float4 m0 = S->mat[3 * mid + 0];
                    const float4 m1 = S->mat[3 * mid + 1];
                    const f3 albedo = xyz(m0);
                    const float metallic = m1.w;
                    const f3 base = add(scale(xyz(m1), 1.0f - metallic), scale(albedo, metallic));
                    const f3 f = cook_torrance(albedo, base, metallic, m0.w, nrm, neg(rd), L);
                    const float4 l0 = S->light[2 * lidx + 0];
                    const float4 l1 = S->light[2 * lidx + 1];
                    col = add(col, mul(mul(scale(xyz(l1), l0.w), f), tp));
                }
                end_bounce = true;
            }
        }
        if (end_bounce) {
            const float4 m0 = S->mat[3 * mid + 0];
            const float4 m1 = S->mat[3 * mid + 1];
            tp = mul(tp, xyz(m0));
            ro = so;
            const float p = gmax(0.1f, gmin(1.0f, length(tp)));
            if (pcg_float(seed) > p) {
                end_path = true;
            } else {
                tp = divs(tp, p);
                rd = m1.w > 0.0f ? sample_ggx(nrm, m0.w, seed) : sample_cosine(nrm, seed);
                ++bounce;
                if (bounce >= S->max_bounces) {
                    end_path = true;
                } else {
                    seed += (uint32_t)bounce;
                    state = kCont;
                }
            }
        }
        if (end_path) {
            // accumulate in frame order (Renderer.cu:164-165) into the chunk's LDS slot
            const SlotView v = slot_view(wave_lds, lslot, P, K);
            float4 acc = v.acc[li];
            if (S->kahan) {
                float4 cp = v.comp[li];
                float yv, t;
                yv = col.x - cp.x; t = acc.x + yv; cp.x = (t - acc.x) - yv; acc.x = t;
                yv = col.y - cp.y; t = acc.y + yv; cp.y = (t - acc.y) - yv; acc.y = t;
                yv = col.z - cp.z; t = acc.z + yv; cp.z = (t - acc.z) - yv; acc.z = t;
                yv = 1.0f - cp.w; t = acc.w + yv; cp.w = (t - acc.w) - yv; acc.w = t;
                v.comp[li] = cp;
            } else {
                acc.x = acc.x + col.x;
                acc.y = acc.y + col.y;
                acc.z = acc.z + col.z;
                acc.w = acc.w + 1.0f;
            }
            v.acc[li] = acc;
            v.done[li] = frame;
            if (S->dbg_radiance)
                S->dbg_radiance[px] = make_float4(col.x, col.y, col.z, 1.0f);
            state = kIdle;
        }
        __builtin_amdgcn_wave_barrier();
    }

    const unsigned long long r = wave_sum(rays);
    const unsigned long long e = wave_sum(exact);
    const unsigned long long ts = wave_sum(tests);
    const unsigned long long bx = wave_sum(boxes);
    if (lane == 0) {
        if (r) atomicAdd(a.counters + kCtrRays, r);
        if (e) atomicAdd(a.counters + kCtrExact, e);
        if (ts) atomicAdd(a.counters + kCtrTests, ts);
        if (bx) atomicAdd(a.counters + kCtrBoxes, bx);
    }
}

__global__ void k_zero_u32(unsigned int* p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

}  // namespace

size_t wavefront_lds_bytes(uint32_t chunk_px, bool kahan)
{
    return 4u * 2u * (size_t)chunk_px * (20u + (kahan ? 16u : 0u));
}

hipError_t launch_wavefront(const WfPass& p, const WfLaunch& l, hipStream_t stream)
{
    if (p.total_chunks == 0)
        return hipSuccess;
    if (p.chunk_px == 0 || p.chunk_px % 64 != 0 || p.chunk_px > kMaxChunkPx)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_render_wavefront, dim3(l.workgroups), dim3(256),
                       wavefront_lds_bytes(p.chunk_px, p.kahan != 0), stream, p);
    return hipGetLastError();
}

hipError_t launch_zero_u32(unsigned int* p, uint32_t n, hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

int wavefront_max_blocks_per_cu(uint32_t chunk_px, bool kahan)
{
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_render_wavefront, 256,
                                                     wavefront_lds_bytes(chunk_px, kahan)) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

}  // namespace atx
