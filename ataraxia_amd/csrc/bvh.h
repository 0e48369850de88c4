// bvh.h -- bounding-volume hierarchy over a scene's spheres: device layout
// (the host builder is bvh_build.h).
//
// The tree only decides which spheres a ray tests.  Every sphere that is
// tested goes through the same filtered test as the flat loop
// (trace_device.h), so the closest hit, including the lowest-index rule for
// equal t (Engine/src/Renderer.cu:274), is the one the reference's loop over
// all spheres finds.  Culling is conservative with respect to the
// reference's rounded arithmetic, not only to exact geometry: see
// trace_bvh.h for the bound that sizes the per-ray box inflation.
//
// Layout in HBM (per scene):
//   nodes  uint4[layouts][n_nodes]   (layouts = 1 or 8), 16 B per node:
//          box corners as 16-bit cells of the root box (world = base +
//          cell * scale per axis), skip (16 bits: node index after this
//          node's subtree; n_nodes = end) and leaf (16 bits: 0 for an inner
//          node, else ((first + 1) << 2) | (count - 1)).
//          Each layout is the same tree in depth-first order; with 8 layouts,
//          layout k visits, at every inner node, first the child nearer along
//          the split axis for ray directions whose sign bits are k (x: 1,
//          y: 2, z: 4); a single layout always visits the lower child first.  The
//          traversal is stackless: a hit inner node continues at index + 1
//          (its first child), anything else at `skip`.
//   lsph   float4[n_tree]   tree spheres in leaf order (cx, cy, cz, r*r)
//   lidx   int32[n_tree]    their index in the scene
//   big    int32[n_big]     spheres kept out of the tree (radius larger than
//                           the region of the others, e.g. a ground sphere,
//                           or non-finite data); tested by a flat loop first
// Boxes are rounded outwards from double precision, so each contains its
// spheres exactly.
#pragma once

#include <stdint.h>
#include <hip/hip_runtime.h>

#include "bvh_build.h"

namespace atx {

struct BvhDev {
    const uint4* nodes;
    const float4* lsph;
    const int32_t* lidx;
    const int32_t* big;
    uint32_t n_nodes;   // nodes per layout; 0 = no tree (flat loop over all spheres)
    uint32_t n_big;
    uint32_t layout_mask;  // 7: one layout per direction octant, 0: one layout
    float cx, cy, cz;   // centre of the tree's region
    float rb;           // upper bound of |centre - sphere centre| over tree spheres
    float r_min, r_max; // lower / upper bound of the tree spheres' radii
    float bx, by, bz;   // world position of cell 0
    float sx, sy, sz;   // world size of one cell
    float isx, isy, isz;  // 1 / size (rounded)
    float ext;          // upper bound of 65535 * cell size over the axes
};

}  // namespace atx
