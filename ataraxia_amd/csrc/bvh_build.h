// bvh_build.h -- host construction of the sphere BVH whose device layout is
// described in bvh.h.  Plain C++ (no HIP types), so that host-side checks can
// build it with any C++ compiler.
#pragma once

#include <stddef.h>
#include <stdint.h>
#include <vector>

namespace atx {

constexpr uint32_t kBvhLayouts = 8;      // maximum number of traversal orders
constexpr uint32_t kBvhLeafMax = 4;      // spheres per leaf (2-bit count field)
constexpr uint32_t kBvhMinSpheres = 16;  // smaller scenes use the flat loop only
constexpr uint32_t kBvhMaxBig = 16;
constexpr uint32_t kBvhMaxTree = 16382;  // tree spheres addressable by a 14-bit leaf index
constexpr uint32_t kBvhGrid = 65535;     // node boxes are 16-bit cells of the tree's box

// Node encoding (4 x u32 = 16 B): cell coordinates of the box corners,
// lo = (x, y, z) and hi = (x, y, z), 16 bits each, then skip (16 bits: node
// index after this node's subtree) and leaf (16 bits: 0 for an inner node,
// else ((first + 1) << 2) | (count - 1)).
struct HostBvh {
    std::vector<uint32_t> nodes; // layouts * n_nodes * 4
    std::vector<float> lsph;     // 4 per tree sphere
    std::vector<int32_t> lidx;
    std::vector<int32_t> big;
    uint32_t n_nodes = 0;
    uint32_t layouts = 0;        // 1 (one depth-first order) or kBvhLayouts (one per direction octant)
    float cx = 0, cy = 0, cz = 0, rb = 0, r_min = 0, r_max = 0;
    float base[3] = {0, 0, 0};   // world position of cell 0
    float scale[3] = {1, 1, 1};  // world size of one cell per axis
    float inv_scale[3] = {1, 1, 1};
    float ext = 0;               // upper bound of kBvhGrid * scale over the axes
};

// Builds the tree over spheres given as (cx, cy, cz, r*r) quadruples, with
// `layouts` traversal orders (1 or kBvhLayouts).  Scenes with fewer than
// kBvhMinSpheres spheres, or more than kBvhMaxBig spheres that cannot go into
// a tree, get no tree (n_nodes = 0).
void build_bvh(const float* sph4, uint32_t n, uint32_t layouts, HostBvh* out);

// Device bytes of the tree of an n-sphere scene (upper bound).
size_t bvh_device_bytes(uint64_t n);

}  // namespace atx
