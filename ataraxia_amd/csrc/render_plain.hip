// render_plain.hip -- the plain single-scene kernel: one lane per pixel, one
// path sample per launch, the reference's per-pixel loop (Renderer.cu:150-170
// kernelRender + Renderer.cu:287-387 perPixel) on gfx950.
//
// Launch shape: 256-lane workgroups cover 16x16 pixels, each wave64 an 8x8
// tile.  The scene's spheres (16 B each) are copied into LDS once per
// workgroup and every closest-hit loop reads them from there: all lanes of a
// wave read the same sphere in the same cycle, so each read is one LDS
// broadcast.  Scenes larger than the LDS budget are read from global memory
// with the same uniform index, which the compiler turns into scalar loads.
//
// Ray generation happens here too (from the inverse matrices the host
// computes once per frame), so no per-pixel direction buffer crosses PCIe.
#include "device_types.h"
#include "trace_bvh.h"
#include "trace_device.h"

namespace atx {
namespace {

struct LdsSrc {
    const float4* p;
    __device__ __forceinline__ float4 operator()(uint32_t i) const { return p[i]; }
};
struct GlobalSrc {
    const float4* __restrict__ p;
    __device__ __forceinline__ float4 operator()(uint32_t i) const { return p[i]; }
};

__device__ __forceinline__ f3 xyz(float4 v) { return mk(v.x, v.y, v.z); }

template <class Src, bool kExactLoop>
__device__ __forceinline__ Hit trace(const Src& src, uint32_t n, f3 o, f3 d, uint32_t& exact_count)
{
    if (kExactLoop)
        return trace_exact(src, n, o, d);
    return trace_filtered(src, n, o, d, exact_count);
}

struct TraceCount {
    uint32_t exact = 0, tests = 0, boxes = 0;
};

// Closest hit (dist2 < 0) or shadow query (dist2 = squared light distance:
// the returned sphere may be any occluder instead of the closest one).
template <bool kLds, bool kExactLoop, bool kBvh>
__device__ __forceinline__ Hit trace_ray(const PlainArgs& a, const LdsSrc& lsrc, const GlobalSrc& gsrc, f3 o, f3 d,
                                         float dist2, TraceCount& c)
{
    if (kBvh) {
        BvhQuery q;
        const Hit h = trace_bvh(a.sc.bvh, gsrc, a.sc.n_sph, o, d, dist2, q);
        c.exact += q.exact;
        c.tests += q.tests;
        c.boxes += q.boxes;
        return h;
    }
    c.tests += a.sc.n_sph;
    return kLds ? trace<LdsSrc, kExactLoop>(lsrc, a.sc.n_sph, o, d, c.exact)
                : trace<GlobalSrc, kExactLoop>(gsrc, a.sc.n_sph, o, d, c.exact);
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

template <bool kLds, bool kKahan, bool kDebug, bool kExactLoop, bool kBvh>
__global__ __launch_bounds__(256) void k_render_plain(PlainArgs a)
{
    extern __shared__ float4 s_sph[];
    if (kLds) {
        for (uint32_t i = threadIdx.x; i < a.sc.n_sph; i += 256)
            s_sph[i] = a.sc.sph[i];
        __syncthreads();
    }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t x = blockIdx.x * 16u + (wave & 1u) * 8u + (lane & 7u);
    const uint32_t y = blockIdx.y * 16u + (wave >> 1) * 8u + (lane >> 3);
    const uint32_t W = a.cam.width, H = a.cam.height;
    uint32_t rays = 0;
    TraceCount cnt;

    if (x < W && y < H) {
        const size_t px = (size_t)x + (size_t)y * W;
        f3 ro = mk(a.cam.pos[0], a.cam.pos[1], a.cam.pos[2]);
        f3 rd = camera_ray(a.cam.inv_proj, a.cam.inv_view, x, y, W, H);
        f3 color = mk(0.0f, 0.0f, 0.0f);
        f3 tp = mk(1.0f, 1.0f, 1.0f);
        uint32_t seed = x + y * W;
        seed *= a.frame;
        const int bounces = a.max_bounces;
        if (kDebug) {
            for (int i = 0; i < (int)a.dbg_bounces; ++i) {
                if (a.dbg_hit) a.dbg_hit[px * a.dbg_bounces + i] = -2;
                if (a.dbg_shadow) a.dbg_shadow[px * a.dbg_bounces + i] = -2;
            }
        }
        const LdsSrc lsrc{s_sph};
        const GlobalSrc gsrc{a.sc.sph};
        for (int i = 0; i < bounces; i++) {
            seed += (uint32_t)i;
            const Hit ht = trace_ray<kLds, kExactLoop, kBvh>(a, lsrc, gsrc, ro, rd, -1.0f, cnt);
            ++rays;
            if (kDebug && a.dbg_hit && i < (int)a.dbg_bounces)
                a.dbg_hit[px * a.dbg_bounces + i] = ht.id;
            if (ht.id < 0) {
                if (a.sky_light)
                    color = add(color, mul(mk(0.6f, 0.7f, 0.9f), tp));
                break;
            }
            // rayHit (Renderer.cu:396-409)
            const float4 sp = a.sc.sph[ht.id];
            const f3 center = xyz(sp);
            f3 pos = add(sub(ro, center), scale(rd, ht.t));
            const f3 nrm = normalize(pos);
            pos = add(pos, center);

            const int32_t m = a.sc.sph_mat[ht.id];
            const float4 m0 = a.sc.mat[3 * m + 0];
            const float4 m1 = a.sc.mat[3 * m + 1];
            const float4 m2 = a.sc.mat[3 * m + 2];
            const f3 albedo = xyz(m0);
            const float roughness = m0.w, metallic = m1.w;
            if (m2.w > 0.0f)
                color = add(color, mul(scale(xyz(m2), m2.w), tp));
            const f3 base = add(scale(xyz(m1), 1.0f - metallic), scale(albedo, metallic));

            if (a.sc.n_light > 0) {
                const uint32_t li = pcg_hash(seed) % a.sc.n_light;
                const float4 l0 = a.sc.light[2 * li + 0];
                const float4 l1 = a.sc.light[2 * li + 1];
                f3 L = sub(xyz(l0), pos);
                const float dist2 = dot(L, L);
                L = normalize(L);
                const f3 so = add(pos, scale(nrm, 0.0001f));
                // the debug capture records the closest occluder; otherwise any proven one will do
                const Hit sh = trace_ray<kLds, kExactLoop, kBvh>(a, lsrc, gsrc, so, L, kDebug ? -1.0f : dist2, cnt);
                ++rays;
                if (kDebug && a.dbg_shadow && i < (int)a.dbg_bounces)
                    a.dbg_shadow[px * a.dbg_bounces + i] = sh.id;
                const float st = sh.id < 0 ? -1.0f : sh.t;
                if (!(st > 0.0f && st * st < dist2)) {
                    const f3 f = cook_torrance(albedo, base, metallic, roughness, nrm, neg(rd), L);
                    const f3 le = scale(xyz(l1), l0.w);
                    color = add(color, mul(mul(le, f), tp));  // pdf = 1 (Renderer.cu:365-367)
                }
            }

            tp = mul(tp, albedo);
            ro = add(pos, scale(nrm, 0.0001f));
            const float p = gmax(0.1f, gmin(1.0f, length(tp)));
            if (pcg_float(seed) > p)
                break;
            tp = divs(tp, p);
            rd = sample_next(nrm, metallic > 0.0f, roughness, seed);
        }

        // kernelRender accumulation and display (Renderer.cu:164-168).  Frame 1
        // starts from zero, which is what the reference's memset provides.
        const float4 c4 = make_float4(color.x, color.y, color.z, 1.0f);
        float4 acc = a.frame == 1u ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : a.accum[px];
        if (kKahan) {
            float4 cp = a.frame == 1u ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : a.comp[px];
            float yv, t;
            yv = c4.x - cp.x; t = acc.x + yv; cp.x = (t - acc.x) - yv; acc.x = t;
            yv = c4.y - cp.y; t = acc.y + yv; cp.y = (t - acc.y) - yv; acc.y = t;
            yv = c4.z - cp.z; t = acc.z + yv; cp.z = (t - acc.z) - yv; acc.z = t;
            yv = c4.w - cp.w; t = acc.w + yv; cp.w = (t - acc.w) - yv; acc.w = t;
            a.comp[px] = cp;
        } else {
            acc.x = acc.x + c4.x;
            acc.y = acc.y + c4.y;
            acc.z = acc.z + c4.z;
            acc.w = acc.w + c4.w;
        }
        a.accum[px] = acc;
        const float fi = (float)a.frame;
        a.image[px] = rgba8(gclamp(acc.x / fi, 0.0f, 1.0f), gclamp(acc.y / fi, 0.0f, 1.0f),
                            gclamp(acc.z / fi, 0.0f, 1.0f), gclamp(acc.w / fi, 0.0f, 1.0f));
        if (kDebug && a.dbg_radiance)
            a.dbg_radiance[px] = c4;
    }

    const unsigned long long r = wave_sum(rays);
    const unsigned long long e = wave_sum(cnt.exact);
    const unsigned long long ts = wave_sum(cnt.tests);
    const unsigned long long bx = wave_sum(cnt.boxes);
    if (lane == 0) {
        if (r) atomicAdd(a.counters + kCtrRays, r);
        if (e) atomicAdd(a.counters + kCtrExact, e);
        if (ts) atomicAdd(a.counters + kCtrTests, ts);
        if (bx) atomicAdd(a.counters + kCtrBoxes, bx);
    }
}

template <bool kLds, bool kKahan, bool kDebug, bool kExactLoop, bool kBvh = false>
hipError_t launch_t(const PlainArgs& a, hipStream_t stream)
{
    const dim3 grid((a.cam.width + 15u) / 16u, (a.cam.height + 15u) / 16u);
    const size_t lds = kLds ? (size_t)a.sc.n_sph * sizeof(float4) : 0;
    hipLaunchKernelGGL((k_render_plain<kLds, kKahan, kDebug, kExactLoop, kBvh>), grid, dim3(256), lds, stream, a);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_plain(const PlainArgs& a, uint32_t v, hipStream_t stream)
{
    const bool lds = v & kVarLds, kahan = v & kVarKahan, dbg = v & kVarDebug, ex = v & kVarExactLoop;
    if (v & kVarBvh) {
        if (a.sc.bvh.n_nodes == 0) return hipErrorInvalidValue;
        if (!kahan && !dbg) return launch_t<false, false, false, false, true>(a, stream);
        if (kahan && !dbg) return launch_t<false, true, false, false, true>(a, stream);
        if (!kahan && dbg) return launch_t<false, false, true, false, true>(a, stream);
        return launch_t<false, true, true, false, true>(a, stream);
    }
#define ATX_DISPATCH(L, K, D, E) \
    if (lds == L && kahan == K && dbg == D && ex == E) return launch_t<L, K, D, E>(a, stream);
    ATX_DISPATCH(true, false, false, false)
    ATX_DISPATCH(true, true, false, false)
    ATX_DISPATCH(true, false, true, false)
    ATX_DISPATCH(true, true, true, false)
    ATX_DISPATCH(false, false, false, false)
    ATX_DISPATCH(false, true, false, false)
    ATX_DISPATCH(false, false, true, false)
    ATX_DISPATCH(false, true, true, false)
    ATX_DISPATCH(true, false, false, true)
    ATX_DISPATCH(true, true, false, true)
    ATX_DISPATCH(true, false, true, true)
    ATX_DISPATCH(true, true, true, true)
    ATX_DISPATCH(false, false, false, true)
    ATX_DISPATCH(false, true, false, true)
    ATX_DISPATCH(false, false, true, true)
    ATX_DISPATCH(false, true, true, true)
#undef ATX_DISPATCH
    return hipErrorInvalidValue;
}

}  // namespace atx
