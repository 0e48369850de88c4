// device_types.h -- device-side scene layout and kernel argument blocks.
//
// HBM layout of one scene (private to libatx; the C ABI takes the reference's
// packed AoS structs and this is built from them on upload):
//   sph      float4[N]   (cx, cy, cz, r*r)      16 B/sphere, the only data the
//                                                closest-hit loop reads
//   sph_mat  int32[N]    material index (clamped to [0, M) on the host)
//   mat      float4[3M]  (albedo, roughness) (F0, metallic)
//                        (emission colour, emission intensity)
//   light    float4[2L]  (position, intensity) (colour, 0)
// r*r is precomputed once; the reference evaluates radius*radius inside its
// loop (Renderer.cu:264), which rounds to the same value.
#pragma once

#include <stdint.h>
#include <hip/hip_runtime.h>

#include "bvh.h"

namespace atx {

struct SceneDev {
    const float4* sph;
    const int32_t* sph_mat;
    const float4* mat;
    const float4* light;
    uint32_t n_sph;
    uint32_t n_mat;
    uint32_t n_light;
    uint32_t pad_;
    BvhDev bvh;              // n_nodes = 0: flat loop over all spheres
};

struct CameraDev {
    float inv_proj[16];
    float inv_view[16];
    float pos[3];
    uint32_t width;
    uint32_t height;
};

// Counters written by the kernels (device memory, u64 each).
enum CounterSlot : uint32_t {
    kCtrRays = 0,          // rays traced (continuation/primary + shadow)
    kCtrExact = 1,         // traces resolved by the exact fallback loop
    kCtrTests = 2,         // ray-sphere tests
    kCtrBoxes = 3,         // ray-box tests of BVH nodes
    kCtrReused = 4,        // rays answered from the per-pixel first-bounce cache
    kCtrCount = 5,
};

struct PlainArgs {
    SceneDev sc;
    CameraDev cam;
    int32_t max_bounces;
    uint32_t sky_light;
    uint32_t frame;          // reference frameIndex, starts at 1
    uint32_t pad_;
    float4* accum;           // width*height
    float4* comp;            // Kahan compensation (nullptr for naive)
    uint32_t* image;         // width*height RGBA8
    float4* dbg_radiance;    // optional: this frame's sample per pixel
    int32_t* dbg_hit;        // optional: [pixel][bounce] continuation hit id
    int32_t* dbg_shadow;     // optional: [pixel][bounce] shadow hit id
    uint32_t dbg_bounces;
    uint32_t pad2_;
    unsigned long long* counters;
};

// Kernel variants.
enum PlainVariant : uint32_t {
    kVarLds = 1u << 0,       // stage the scene's spheres in LDS
    kVarKahan = 1u << 1,     // compensated accumulation
    kVarDebug = 1u << 2,     // write per-pixel debug capture
    kVarExactLoop = 1u << 3, // closest-hit loop with exact sqrt/div per sphere
    kVarBvh = 1u << 4,       // BVH traversal (sc.bvh); excludes kVarLds and kVarExactLoop
};

hipError_t launch_plain(const PlainArgs& a, uint32_t variant, hipStream_t stream);

}  // namespace atx
