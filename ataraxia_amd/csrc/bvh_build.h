// bvh_build.h -- host construction of the sphere BVH whose device layout is
// described in bvh.h.  Plain C++ (no HIP types), so that host-side checks can
// build it with any C++ compiler.
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <vector>

namespace atx {

constexpr uint32_t kBvhLayouts = 8;      // maximum number of traversal orders
constexpr uint32_t kBvhLeafMax = 4;      // spheres per leaf (3-bit count field)
constexpr uint32_t kBvhMinSpheres = 16;  // smaller scenes use the flat loop only
constexpr uint32_t kBvhMaxBig = 16;

struct HostBvh {
    std::vector<float> nodes;    // layouts * n_nodes * 8
    std::vector<float> lsph;     // 4 per tree sphere
    std::vector<int32_t> lidx;
    std::vector<int32_t> big;
    uint32_t n_nodes = 0;
    uint32_t layouts = 0;        // 1 (one depth-first order) or kBvhLayouts (one per direction octant)
    float cx = 0, cy = 0, cz = 0, rb = 0, r_min = 0, r_max = 0, cbox = 0;
};

// Builds the tree over spheres given as (cx, cy, cz, r*r) quadruples, with
// `layouts` traversal orders (1 or kBvhLayouts).  Scenes with fewer than
// kBvhMinSpheres spheres, or more than kBvhMaxBig spheres that cannot go into
// a tree, get no tree (n_nodes = 0).
void build_bvh(const float* sph4, uint32_t n, uint32_t layouts, HostBvh* out);

// Device bytes of the tree of an n-sphere scene (upper bound).
size_t bvh_device_bytes(uint64_t n);

}  // namespace atx
