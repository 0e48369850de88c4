// bvh_build.cpp -- host construction of the sphere BVH (device layout in bvh.h).
//
// Binary tree, surface-area heuristic over sphere boxes, at most kBvhLeafMax
// spheres per leaf.  All geometry is evaluated in double precision; every box
// is rounded outwards to float and then to 16-bit cells of the root box (one
// extra cell on each side), so a node box contains each of its spheres
// (centre c, radius sqrt(r*r)) exactly.  Spheres whose radius exceeds
// the half-diagonal of the region of all other centres (a ground sphere of
// radius 1000 under a cluster of unit spheres) and spheres with non-finite or
// negative data are kept out of the tree: they would make every box, and
// the per-ray inflation that scales with the region, as large as the scene.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "bvh_build.h"

namespace atx {
namespace {

struct Prim {
    double lo[3], hi[3], c[3];
    int32_t idx;
};

struct BNode {
    float lo[3], hi[3];
    int32_t left = -1, right = -1;
    int32_t axis = 0;
    uint32_t first = 0, count = 0;
};

float round_down(double x)
{
    float f = (float)x;
    if ((double)f > x) f = std::nextafter(f, -INFINITY);
    return f;
}

float round_up(double x)
{
    float f = (float)x;
    if ((double)f < x) f = std::nextafter(f, INFINITY);
    return f;
}

double half_area(const double lo[3], const double hi[3])
{
    const double dx = std::max(0.0, hi[0] - lo[0]), dy = std::max(0.0, hi[1] - lo[1]),
                 dz = std::max(0.0, hi[2] - lo[2]);
    return dx * dy + dy * dz + dz * dx;
}

struct Builder {
    std::vector<Prim>& P;
    std::vector<BNode> nodes;
    std::vector<Prim> leaf_order;
    static constexpr double kCostBox = 0.5;   // box test relative to one sphere test

    explicit Builder(std::vector<Prim>& p) : P(p) {}

    void bounds(int b, int e, double lo[3], double hi[3]) const
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = INFINITY;
            hi[k] = -INFINITY;
        }
        for (int i = b; i < e; ++i)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], P[i].lo[k]);
                hi[k] = std::max(hi[k], P[i].hi[k]);
            }
    }

    void sort_axis(int b, int e, int axis)
    {
        std::sort(P.begin() + b, P.begin() + e, [axis](const Prim& x, const Prim& y) {
            return x.c[axis] < y.c[axis] || (x.c[axis] == y.c[axis] && x.idx < y.idx);
        });
    }

    int build(int b, int e)
    {
        const int id = (int)nodes.size();
        nodes.emplace_back();
        double lo[3], hi[3];
        bounds(b, e, lo, hi);
        for (int k = 0; k < 3; ++k) {
            nodes[id].lo[k] = round_down(lo[k]);
            nodes[id].hi[k] = round_up(hi[k]);
        }
        const int n = e - b;
        const double parent = half_area(lo, hi);
        double best_cost = INFINITY;
        int best_axis = -1, best_split = -1;
        std::vector<double> left_area((size_t)n + 1);
        for (int axis = 0; axis < 3 && n > 1; ++axis) {
            sort_axis(b, e, axis);
            double l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int i = 1; i < n; ++i) {
                for (int k = 0; k < 3; ++k) {
                    l[k] = std::min(l[k], P[b + i - 1].lo[k]);
                    h[k] = std::max(h[k], P[b + i - 1].hi[k]);
                }
                left_area[i] = half_area(l, h);
            }
            double rl[3] = {INFINITY, INFINITY, INFINITY}, rh[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (int i = n - 1; i >= 1; --i) {
                for (int k = 0; k < 3; ++k) {
                    rl[k] = std::min(rl[k], P[b + i].lo[k]);
                    rh[k] = std::max(rh[k], P[b + i].hi[k]);
                }
                const double cost = left_area[i] * i + half_area(rl, rh) * (n - i);
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = i;
                }
            }
        }
        const double split_cost =
            parent > 0.0 ? kCostBox + best_cost / parent : kCostBox + (double)n;
        const bool leaf = n == 1 || (n <= (int)kBvhLeafMax && (double)n <= split_cost);
        if (leaf || best_axis < 0) {
            nodes[id].first = (uint32_t)leaf_order.size();
            nodes[id].count = (uint32_t)n;
            for (int i = b; i < e; ++i)
                leaf_order.push_back(P[i]);
            return id;
        }
        sort_axis(b, e, best_axis);
        nodes[id].axis = best_axis;
        const int left = build(b, b + best_split);
        const int right = build(b + best_split, e);
        nodes[id].left = left;
        nodes[id].right = right;
        return id;
    }
};

struct Grid {
    double base[3], scale[3];
};

uint32_t cell_down(double x, const Grid& g, int a)
{
    const double c = std::floor((x - g.base[a]) / g.scale[a]) - 1.0;
    return (uint32_t)std::min(std::max(c, 0.0), (double)kBvhGrid);
}

uint32_t cell_up(double x, const Grid& g, int a)
{
    const double c = std::ceil((x - g.base[a]) / g.scale[a]) + 1.0;
    return (uint32_t)std::min(std::max(c, 0.0), (double)kBvhGrid);
}

void emit(const std::vector<BNode>& T, int id, uint32_t oct, const Grid& g, std::vector<uint32_t>& out,
          uint32_t& cur)
{
    const uint32_t me = cur++;
    uint32_t* d = out.data() + (size_t)me * 4;
    const BNode& nd = T[id];
    uint32_t skip, leaf;
    if (nd.left < 0) {
        skip = me + 1;
        leaf = ((nd.first + 1u) << 2) | (nd.count - 1u);
    } else {
        const bool neg = (oct >> nd.axis) & 1u;
        emit(T, neg ? nd.right : nd.left, oct, g, out, cur);
        emit(T, neg ? nd.left : nd.right, oct, g, out, cur);
        skip = cur;
        leaf = 0;
    }
    const uint32_t lx = cell_down(nd.lo[0], g, 0), ly = cell_down(nd.lo[1], g, 1), lz = cell_down(nd.lo[2], g, 2);
    const uint32_t hx = cell_up(nd.hi[0], g, 0), hy = cell_up(nd.hi[1], g, 1), hz = cell_up(nd.hi[2], g, 2);
    d[0] = lx | (ly << 16);
    d[1] = lz | (hx << 16);
    d[2] = hy | (hz << 16);
    d[3] = skip | (leaf << 16);
}

bool finite_sphere(const float* s)
{
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(s[k])) return false;
    return s[3] >= 0.0f && std::fabs(s[0]) < 1e15f && std::fabs(s[1]) < 1e15f && std::fabs(s[2]) < 1e15f &&
           s[3] < 1e30f;
}

}  // namespace

void build_bvh(const float* sph4, uint32_t n, uint32_t layouts, HostBvh* out)
{
    *out = HostBvh();
    layouts = layouts >= kBvhLayouts ? kBvhLayouts : 1u;
    if (n < kBvhMinSpheres) return;

    std::vector<int32_t> big;
    std::vector<char> is_big(n, 0);
    std::vector<uint32_t> valid;
    for (uint32_t i = 0; i < n; ++i) {
        if (finite_sphere(sph4 + 4 * i)) {
            valid.push_back(i);
        } else {
            big.push_back((int32_t)i);
            is_big[i] = 1;
        }
    }
    // Largest spheres first: a sphere goes out of the tree while its radius
    // exceeds the half-diagonal of the box around all remaining centres.
    std::vector<uint32_t> by_r = valid;
    std::sort(by_r.begin(), by_r.end(), [&](uint32_t a, uint32_t b) {
        return sph4[4 * a + 3] > sph4[4 * b + 3] || (sph4[4 * a + 3] == sph4[4 * b + 3] && a < b);
    });
    for (size_t k = 0; k < by_r.size() && big.size() < kBvhMaxBig + 1; ++k) {
        const uint32_t cand = by_r[k];
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i : valid) {
            if (is_big[i] || i == cand) continue;
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], (double)sph4[4 * i + a]);
                hi[a] = std::max(hi[a], (double)sph4[4 * i + a]);
            }
        }
        if (!(lo[0] <= hi[0])) break;
        const double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        const double half_diag = 0.5 * std::sqrt(dx * dx + dy * dy + dz * dz);
        if (std::sqrt((double)sph4[4 * cand + 3]) > half_diag) {
            big.push_back((int32_t)cand);
            is_big[cand] = 1;
        } else {
            break;
        }
    }
    std::vector<Prim> P;
    for (uint32_t i : valid) {
        if (is_big[i]) continue;
        Prim p;
        const double r = std::sqrt((double)sph4[4 * i + 3]) * (1.0 + 0x1p-40);
        for (int a = 0; a < 3; ++a) {
            p.c[a] = sph4[4 * i + a];
            p.lo[a] = p.c[a] - r;
            p.hi[a] = p.c[a] + r;
        }
        p.idx = (int32_t)i;
        P.push_back(p);
    }
    if (big.size() > kBvhMaxBig || P.size() < kBvhMinSpheres || P.size() > kBvhMaxTree) return;
    std::sort(big.begin(), big.end());

    // Region of the tree: centre of the centres' box, bounds of distance and radius.
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double r2_min = INFINITY, r2_max = 0.0;
    for (const Prim& p : P) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], p.c[a]);
            hi[a] = std::max(hi[a], p.c[a]);
        }
        const double r2 = sph4[4 * p.idx + 3];
        r2_min = std::min(r2_min, r2);
        r2_max = std::max(r2_max, r2);
    }
    const double C[3] = {0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
    const float Cf[3] = {(float)C[0], (float)C[1], (float)C[2]};
    double rb = 0.0;
    for (const Prim& p : P) {
        const double dx = p.c[0] - Cf[0], dy = p.c[1] - Cf[1], dz = p.c[2] - Cf[2];
        rb = std::max(rb, std::sqrt(dx * dx + dy * dy + dz * dz));
    }

    Builder B(P);
    B.build(0, (int)P.size());
    const uint32_t nn = (uint32_t)B.nodes.size();

    // Cell grid over the root box: world = base + cell * scale, with scale a
    // float so that the device's grid coordinates mean exactly the same boxes.
    // Boxes round outwards and get one extra cell on each side.
    Grid g;
    for (int a = 0; a < 3; ++a) {
        const double lo_a = B.nodes[0].lo[a], hi_a = B.nodes[0].hi[a];
        const double span = hi_a - lo_a;
        float sc = round_up(span > 0.0 ? span / (double)(kBvhGrid - 4) : 1.0);
        if (!(sc > 0.0f)) sc = 1.0f;
        g.base[a] = lo_a;
        g.scale[a] = sc;
        out->base[a] = B.nodes[0].lo[a];
        out->scale[a] = sc;
        out->inv_scale[a] = (float)(1.0 / (double)sc);
    }
    out->ext = round_up((double)kBvhGrid * std::max(std::max(g.scale[0], g.scale[1]), g.scale[2]) * (1.0 + 0x1p-30));
    out->n_nodes = nn;
    out->layouts = layouts;
    out->nodes.assign((size_t)layouts * nn * 4, 0u);
    for (uint32_t oct = 0; oct < layouts; ++oct) {
        std::vector<uint32_t> lay((size_t)nn * 4);
        uint32_t cur = 0;
        emit(B.nodes, 0, oct, g, lay, cur);
        std::copy(lay.begin(), lay.end(), out->nodes.begin() + (size_t)oct * nn * 4);
    }
    for (const Prim& p : B.leaf_order) {
        for (int a = 0; a < 4; ++a)
            out->lsph.push_back(sph4[4 * p.idx + a]);
        out->lidx.push_back(p.idx);
    }
    out->big = big;
    out->cx = Cf[0];
    out->cy = Cf[1];
    out->cz = Cf[2];
    out->rb = round_up(rb * (1.0 + 0x1p-30));
    out->r_min = round_down(std::sqrt(r2_min) * (1.0 - 0x1p-30));
    out->r_max = round_up(std::sqrt(r2_max) * (1.0 + 0x1p-30));
}

size_t bvh_device_bytes(uint64_t n)
{
    const uint64_t nodes = n > 0 ? 2 * n - 1 : 0;
    return (size_t)(kBvhLayouts * nodes * 16 + 16) + (size_t)(16 * n + 16) + (size_t)(4 * n + 16) +
           (size_t)(4 * (kBvhMaxBig + 1) + 16);
}

}  // namespace atx
