// bvh_check.cpp -- host-side check of the sphere BVH (built by
// ataraxia_amd/csrc/bvh_build.cpp) against the oracle's flat closest-hit loop
// (oracle/atx_oracle.c, ora_trace_ray: the reference's Renderer.cu:251-285).
//
// 1. Structure: every layout is a depth-first order of one tree with correct
//    skip links; child cell boxes nest in parent boxes; every tree sphere's
//    exact box lies in its leaf's cell box; leaves plus the out-of-tree list
//    cover each sphere exactly once.
// 2. Culling: a float restatement of the traversal in trace_bvh.h (with
//    correctly rounded sqrt and reciprocal where the device uses the
//    hardware's approximations) must return the oracle's closest hit (index
//    and t bit for bit) and the oracle's shadow decision (t > 0 && t*t <
//    dist2) for rays built to stress it: grazing rays aimed at silhouettes,
//    duplicate spheres (ties), axis-parallel and signed-zero directions,
//    origins inside spheres and far outside the scene.
// Prints one line per scene; exits non-zero on the first mismatch.
//
// Build: g++ -std=c++17 -O2 -ffp-contract=off bvh_check.cpp bvh_build.cpp atx_oracle.c
#include <cfloat>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../ataraxia_amd/csrc/bvh_build.h"
extern "C" {
#include "../oracle/atx_oracle.h"
}

using atx::HostBvh;

namespace {

struct Rng {  // splitmix64
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float uni(float lo, float hi) { return lo + (hi - lo) * (float)((next() >> 40) * 0x1p-24); }
};

struct V { float x, y, z; };
V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V vadd(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V vscale(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float vdot(V a, V b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
V vnorm(V a) { return vscale(a, 1.0f / std::sqrt(vdot(a, a))); }

// ---- restatement of trace_device.h / trace_bvh.h ---------------------------
struct Filter {
    float lo = INFINITY, hi = INFINITY;
    int32_t best = -1;
    bool amb = false;
};

bool filter_sphere(const float* s, int32_t id, V o, V d, float a, Filter& f)
{
    const float ocx = o.x - s[0], ocy = o.y - s[1], ocz = o.z - s[2];
    const float b = (ocx * d.x + ocy * d.y) + ocz * d.z;
    const float c = ((ocx * ocx + ocy * ocy) + ocz * ocz) - s[3];
    const float disc = b * b - a * c;
    const float sq = std::sqrt(disc);
    const float nn = -b - sq;
    const float u = std::fmax(sq + std::fabs(nn), 0x1p-30f);
    const float hi = std::fma(u, 0x1p-17f, nn);
    const float lo = std::fma(u, -0x1p-17f, nn);
    const bool better = (hi < f.lo) && (lo > 0.0f);
    const bool near = (hi > 0.0f) && (lo < f.hi);
    f.amb = f.amb || (near && !better);
    if (better) {
        f.lo = lo;
        f.hi = hi;
        f.best = id;
    }
    return better;
}

float exact_t(const float* s, V o, V d, float a)
{
    const float ocx = o.x - s[0], ocy = o.y - s[1], ocz = o.z - s[2];
    const float b = (ocx * d.x + ocy * d.y) + ocz * d.z;
    const float c = ((ocx * ocx + ocy * ocy) + ocz * ocz) - s[3];
    const float disc = b * b - a * c;
    return (-b - std::sqrt(disc)) / a;
}

float inv_dir(float v)
{
    const float c = std::fabs(v) < 0x1p-64f ? std::copysign(0x1p-64f, v) : v;
    return 1.0f / c;
}

float excess(float r, float D2)
{
    const float q = 0x1p-17f * (D2 + r * r);
    return q * (1.0f / (std::sqrt(r * r + q) + r));
}

struct Query {
    uint64_t tests = 0, boxes = 0, exact = 0;
};

// returns sphere index (-1 miss); *t gets its exact t
int32_t trace_bvh(const HostBvh& B, const std::vector<float>& sph, V o, V d, float dist2, float* t, Query& q)
{
    const uint32_t n_all = (uint32_t)(sph.size() / 4);
    const float a = vdot(d, d);
    const bool shadow = dist2 >= 0.0f;
    Filter f;
    for (int32_t id : B.big) filter_sphere(&sph[4 * id], id, o, d, a, f);
    q.tests += B.big.size();
    const float ox = o.x - B.cx, oy = o.y - B.cy, oz = o.z - B.cz;
    const float dist = std::sqrt((ox * ox + oy * oy) + oz * oz);
    const float D = (dist * (1.0f + 0x1p-18f) + B.rb) * (1.0f + 0x1p-20f);
    const float D2 = D * D;
    const float m = std::fmax(excess(B.r_min, D2), excess(B.r_max, D2)) * (1.0f + 0x1p-16f);
    const float rx = o.x - B.base[0], ry = o.y - B.base[1], rz = o.z - B.base[2];
    const float omax = std::fmax(std::fmax(std::fabs(rx), std::fabs(ry)), std::fabs(rz));
    const float e = 0x1p-18f * ((omax + B.ext) + m) + 0x1p-60f;
    const float M = (m + e) * (1.0f + 0x1p-20f);
    bool flat = !(M < 0x1p60f) || !(omax < 0x1p60f);
    if (!flat) {
        const float ix = inv_dir(d.x) * B.scale[0], iy = inv_dir(d.y) * B.scale[1], iz = inv_dir(d.z) * B.scale[2];
        const float gx = rx * B.inv_scale[0], gy = ry * B.inv_scale[1], gz = rz * B.inv_scale[2];
        const float mx = M * B.inv_scale[0] * (1.0f + 0x1p-20f), my = M * B.inv_scale[1] * (1.0f + 0x1p-20f),
                    mz = M * B.inv_scale[2] * (1.0f + 0x1p-20f);
        const float px = -((gx + mx) * ix), py = -((gy + my) * iy), pz = -((gz + mz) * iz);
        const float qx = -((gx - mx) * ix), qy = -((gy - my) * iy), qz = -((gz - mz) * iz);
        const float inv_a_up = (1.0f / a) * (1.0f + 0x1p-20f);
        const float s_up = shadow ? std::sqrt(dist2) * (1.0f + 0x1p-18f) : INFINITY;
        float tlim = std::fmin(f.hi * inv_a_up, s_up);
        const uint32_t oct = ((d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u)) &
                             (B.layouts - 1u);
        const uint32_t* nodes = B.nodes.data() + (size_t)oct * B.n_nodes * 4;
        auto slab = [&](const uint32_t* w, float& tn, float& tf) {
            const float lx = (float)(w[0] & 0xffffu), ly = (float)(w[0] >> 16), lz = (float)(w[1] & 0xffffu);
            const float hx = (float)(w[1] >> 16), hy = (float)(w[2] & 0xffffu), hz = (float)(w[2] >> 16);
            const float ax = std::fma(lx, ix, px), bx = std::fma(hx, ix, qx);
            const float ay = std::fma(ly, iy, py), by = std::fma(hy, iy, qy);
            const float az = std::fma(lz, iz, pz), bz = std::fma(hz, iz, qz);
            tn = std::fmax(std::fmax(std::fmin(ax, bx), std::fmin(ay, by)), std::fmax(std::fmin(az, bz), 0.0f));
            tf = std::fmin(std::fmin(std::fmax(ax, bx), std::fmax(ay, by)), std::fmax(az, bz));
        };
        auto leaf_spheres = [&](uint32_t leaf) {
            const uint32_t first = (leaf >> 2) - 1u, cnt = (leaf & 3u) + 1u;
            for (uint32_t k = 0; k < cnt; ++k)
                if (filter_sphere(&B.lsph[4 * (first + k)], B.lidx[first + k], o, d, a, f))
                    tlim = std::fmin(tlim, f.hi * inv_a_up);
            q.tests += cnt;
        };
        // one node per step, as in trace_bvh.h
        uint32_t i = (shadow && tlim * tlim < dist2) ? B.n_nodes : 0u;
        while (i < B.n_nodes) {
            const uint32_t* w = nodes + (size_t)i * 4;
            ++q.boxes;
            float tn, tf;
            slab(w, tn, tf);
            const bool hit = tn <= std::fmin(tf, tlim);
            const uint32_t leaf = w[3] >> 16;
            uint32_t next = hit ? i + 1 : (w[3] & 0xffffu);   // a leaf's skip is i + 1
            if (hit && leaf != 0) {
                leaf_spheres(leaf);
                if (shadow && tlim * tlim < dist2) next = B.n_nodes;
            }
            i = next;
        }
    } else {
        for (uint32_t k = 0; k < n_all; ++k) filter_sphere(&sph[4 * k], (int32_t)k, o, d, a, f);
        q.tests += n_all;
    }
    if (f.amb) {
        ++q.exact;
        float tm = FLT_MAX;
        int32_t best = -1;
        for (uint32_t k = 0; k < n_all; ++k) {
            const float* s = &sph[4 * k];
            const float ocx = o.x - s[0], ocy = o.y - s[1], ocz = o.z - s[2];
            const float b = (ocx * d.x + ocy * d.y) + ocz * d.z;
            const float c = ((ocx * ocx + ocy * ocy) + ocz * ocz) - s[3];
            const float disc = b * b - a * c;
            if (disc >= 0.0f) {
                const float tt = (-b - std::sqrt(disc)) / a;
                if (tt > 0.0f && tt < tm) {
                    tm = tt;
                    best = (int32_t)k;
                }
            }
        }
        *t = tm;
        return best;
    }
    *t = f.best >= 0 ? exact_t(&sph[4 * f.best], o, d, a) : FLT_MAX;
    return f.best;
}

// ---- structure ---------------------------------------------------------------
int fail(const char* what, const char* scene)
{
    std::fprintf(stderr, "FAIL [%s]: %s\n", scene, what);
    return 1;
}

// Cell box of node word w as world coordinates (double), per the grid.
void node_box(const HostBvh& B, const uint32_t* w, double lo[3], double hi[3])
{
    const uint32_t c[6] = {w[0] & 0xffffu, w[0] >> 16, w[1] & 0xffffu, w[1] >> 16, w[2] & 0xffffu, w[2] >> 16};
    for (int a = 0; a < 3; ++a) {
        lo[a] = (double)B.base[a] + (double)c[a] * (double)B.scale[a];
        hi[a] = (double)B.base[a] + (double)c[3 + a] * (double)B.scale[a];
    }
}

bool box_has_sphere(const double lo[3], const double hi[3], const float* s)
{
    const double r = std::sqrt((double)s[3]);
    for (int a = 0; a < 3; ++a)
        if ((double)s[a] - r < lo[a] || (double)s[a] + r > hi[a]) return false;
    return true;
}

int check_structure(const HostBvh& B, const std::vector<float>& sph, const char* name)
{
    const uint32_t n_all = (uint32_t)(sph.size() / 4);
    std::vector<int> seen(n_all, 0);
    for (int32_t id : B.big) seen[id]++;
    for (int32_t id : B.lidx) seen[id]++;
    for (uint32_t i = 0; i < n_all; ++i)
        if (seen[i] != 1) return fail("sphere not covered exactly once", name);
    for (uint32_t oct = 0; oct < B.layouts; ++oct) {
        const uint32_t* nodes = B.nodes.data() + (size_t)oct * B.n_nodes * 4;
        // walk: a subtree of node i spans [i, skip(i)); its children tile (i, skip)
        uint32_t leaf_spheres = 0;
        for (uint32_t i = 0; i < B.n_nodes; ++i) {
            const uint32_t* n = nodes + (size_t)i * 4;
            const uint32_t skip = n[3] & 0xffffu, leaf = n[3] >> 16;
            if (skip <= i || skip > B.n_nodes) return fail("skip link out of range", name);
            if (leaf != 0 && skip != i + 1) return fail("leaf skip is not the next node", name);
            double lo[3], hi[3];
            node_box(B, n, lo, hi);
            if (i > 0) {
                // parent: the nearest earlier node whose subtree contains i
                uint32_t p = i - 1;
                for (;;) {
                    if ((nodes[(size_t)p * 4 + 3] & 0xffffu) > i) break;
                    if (p == 0) return fail("node outside the root's subtree", name);
                    --p;
                }
                const uint32_t* pn = nodes + (size_t)p * 4;
                if (skip > (pn[3] & 0xffffu)) return fail("subtree overlaps its parent's end", name);
                double plo[3], phi[3];
                node_box(B, pn, plo, phi);
                for (int a = 0; a < 3; ++a)
                    if (lo[a] < plo[a] || hi[a] > phi[a]) return fail("child box outside parent box", name);
            }
            if (leaf != 0) {
                const uint32_t first = (leaf >> 2) - 1u, cnt = (leaf & 3u) + 1u;
                if (cnt > atx::kBvhLeafMax || first + cnt > B.lidx.size()) return fail("bad leaf range", name);
                for (uint32_t k = 0; k < cnt; ++k) {
                    if (!box_has_sphere(lo, hi, &B.lsph[4 * (first + k)])) return fail("sphere outside leaf box", name);
                    if (std::memcmp(&B.lsph[4 * (first + k)], &sph[4 * B.lidx[first + k]], 16) != 0)
                        return fail("leaf sphere differs from scene sphere", name);
                }
                leaf_spheres += cnt;
            }
        }
        if ((nodes[3] & 0xffffu) != B.n_nodes) return fail("root skip is not the end", name);
        if (leaf_spheres != B.lidx.size()) return fail("leaves do not hold every tree sphere", name);
    }
    return 0;
}

// ---- scenes and rays -----------------------------------------------------------
struct Scene {
    const char* name;
    std::vector<ora_sphere> s;
};

void push(Scene& sc, float x, float y, float z, float r)
{
    ora_sphere o{};
    o.center[0] = x;
    o.center[1] = y;
    o.center[2] = z;
    o.radius = r;
    o.material = 0;
    sc.s.push_back(o);
}

std::vector<Scene> make_scenes()
{
    std::vector<Scene> v;
    Rng g{1};
    {
        Scene sc{"random_257_ground", {}};
        push(sc, 0, -1000, 0, 1000);
        for (int i = 0; i < 256; ++i)
            push(sc, g.uni(-20, 20), g.uni(0, 6), g.uni(-20, 20), g.uni(0.3f, 1.5f));
        v.push_back(sc);
    }
    {
        Scene sc{"duplicates_ties", {}};
        for (int i = 0; i < 48; ++i) {
            const float x = g.uni(-5, 5), y = g.uni(-5, 5), z = g.uni(-5, 5), r = g.uni(0.2f, 1.0f);
            push(sc, x, y, z, r);
            push(sc, x, y, z, r);   // same sphere at a later index: ties resolved by index
        }
        v.push_back(sc);
    }
    {
        Scene sc{"tiny_and_large", {}};
        for (int i = 0; i < 200; ++i)
            push(sc, g.uni(-10, 10), g.uni(-10, 10), g.uni(-10, 10), i % 4 == 0 ? 1e-3f : g.uni(0.5f, 3.0f));
        push(sc, 0, 0, 0, 0.0f);
        v.push_back(sc);
    }
    {
        Scene sc{"offset_1e4", {}};
        for (int i = 0; i < 128; ++i)
            push(sc, 1e4f + g.uni(-3, 3), -2e4f + g.uni(-3, 3), 5e3f + g.uni(-3, 3), g.uni(0.05f, 0.6f));
        v.push_back(sc);
    }
    {
        Scene sc{"touching_clusters", {}};
        for (int i = 0; i < 64; ++i) {
            const float x = (float)(i % 8) * 2.0f, z = (float)(i / 8) * 2.0f;
            push(sc, x, 0, z, 1.0f);          // neighbours touch exactly
            push(sc, x + 1.0f, 1.0f, z, 0.5f);
        }
        push(sc, 0, -500, 0, 499.0f);
        push(sc, 0, 500, 0, 400.0f);
        v.push_back(sc);
    }
    return v;
}

V random_dir(Rng& g)
{
    for (;;) {
        V d{g.uni(-1, 1), g.uni(-1, 1), g.uni(-1, 1)};
        const float l2 = vdot(d, d);
        if (l2 > 1e-4f && l2 <= 1.0f) return vnorm(d);
    }
}

}  // namespace

int main(int argc, char** argv)
{
    const int rays_per_kind = argc > 1 ? std::atoi(argv[1]) : 20000;
    const uint32_t layouts = argc > 2 ? (uint32_t)std::atoi(argv[2]) : atx::kBvhLayouts;
    int rc = 0;
    for (Scene& sc : make_scenes()) {
        const uint32_t n = (uint32_t)sc.s.size();
        std::vector<float> sph(4 * (size_t)n);
        for (uint32_t i = 0; i < n; ++i) {
            sph[4 * i + 0] = sc.s[i].center[0];
            sph[4 * i + 1] = sc.s[i].center[1];
            sph[4 * i + 2] = sc.s[i].center[2];
            sph[4 * i + 3] = sc.s[i].radius * sc.s[i].radius;
        }
        HostBvh B;
        atx::build_bvh(sph.data(), n, layouts, &B);
        if (B.n_nodes == 0) return fail("no tree built", sc.name);
        if (check_structure(B, sph, sc.name)) return 1;

        Rng g{77 + n};
        Query q;
        uint64_t nrays = 0, hits = 0, occluded = 0;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = 0; i < n; ++i)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::fmin(lo[a], sph[4 * i + a]);
                hi[a] = std::fmax(hi[a], sph[4 * i + a]);
            }
        auto pick_origin = [&](int kind) -> V {
            const uint32_t s = (uint32_t)(g.next() % n);
            const float* c = &sph[4 * s];
            const float r = std::sqrt(c[3]);
            switch (kind) {
            case 0:   // in the scene's box, slightly enlarged
                return {g.uni(lo[0] - 2, hi[0] + 2), g.uni(lo[1] - 2, hi[1] + 2), g.uni(lo[2] - 2, hi[2] + 2)};
            case 1: {  // just outside a sphere's surface, as shadow/continuation origins are
                const V nrm = random_dir(g);
                const V p = vadd(V{c[0], c[1], c[2]}, vscale(nrm, r));
                return vadd(p, vscale(nrm, 1e-4f));
            }
            case 2:   // inside a sphere
                return {c[0] + 0.3f * r, c[1] - 0.2f * r, c[2]};
            default: {  // far away (camera-like, 4x the scene extent)
                const V dir = random_dir(g);
                const float ext = (hi[0] - lo[0]) + (hi[1] - lo[1]) + (hi[2] - lo[2]) + 1.0f;
                return vadd(V{0.5f * (lo[0] + hi[0]), 0.5f * (lo[1] + hi[1]), 0.5f * (lo[2] + hi[2])},
                            vscale(dir, 4.0f * ext));
            }
            }
        };
        auto dir_for = [&](V o, int kind) -> V {
            if (kind == 0) return random_dir(g);
            if (kind == 1) {   // grazing: aim at a point on a sphere's silhouette
                const uint32_t s = (uint32_t)(g.next() % n);
                const float* c = &sph[4 * s];
                const V toc = vsub(V{c[0], c[1], c[2]}, o);
                V perp{toc.y, -toc.x, 0.0f};
                if (vdot(perp, perp) < 1e-12f) perp = {0.0f, toc.z, -toc.y};
                perp = vnorm(perp);
                const float jitter = 1.0f + g.uni(-4e-7f, 4e-7f);
                const V target = vadd(V{c[0], c[1], c[2]}, vscale(perp, std::sqrt(c[3]) * jitter));
                const V d = vsub(target, o);
                return vdot(d, d) > 0.0f ? vnorm(d) : random_dir(g);
            }
            if (kind == 2) {   // axis-parallel with signed zeros
                const int axis = (int)(g.next() % 3);
                const float sgn = (g.next() & 1) ? 1.0f : -1.0f;
                V d{(g.next() & 1) ? 0.0f : -0.0f, (g.next() & 1) ? 0.0f : -0.0f, (g.next() & 1) ? 0.0f : -0.0f};
                (axis == 0 ? d.x : axis == 1 ? d.y : d.z) = sgn;
                return d;
            }
            // towards a sphere centre
            const uint32_t s = (uint32_t)(g.next() % n);
            const V d = vsub(V{sph[4 * s], sph[4 * s + 1], sph[4 * s + 2]}, o);
            return vdot(d, d) > 0.0f ? vnorm(d) : random_dir(g);
        };
        for (int ok = 0; ok < 4; ++ok)
            for (int dk = 0; dk < 4; ++dk)
                for (int r = 0; r < rays_per_kind / 16; ++r) {
                    const V o = pick_origin(ok);
                    const V d = dir_for(o, dk);
                    const float o3[3] = {o.x, o.y, o.z}, d3[3] = {d.x, d.y, d.z};
                    float tr = 0.0f, tb = 0.0f;
                    const int32_t ir = ora_trace_ray(o3, d3, sc.s.data(), n, &tr);
                    const int32_t ib = trace_bvh(B, sph, o, d, -1.0f, &tb, q);
                    ++nrays;
                    if (ir != ib || std::memcmp(&tr, &tb, 4) != 0) {
                        std::fprintf(stderr,
                                     "FAIL [%s] closest: o=(%a,%a,%a) d=(%a,%a,%a) oracle=%d t=%a bvh=%d t=%a\n",
                                     sc.name, o.x, o.y, o.z, d.x, d.y, d.z, ir, tr, ib, tb);
                        return 1;
                    }
                    hits += ir >= 0;
                    // shadow query towards a point along the ray
                    const float along = g.uni(0.0f, 2.0f) * (ir >= 0 ? tr : 10.0f);
                    const V lp = vadd(o, vscale(d, along));
                    V L = vsub(lp, o);
                    const float dist2 = vdot(L, L);
                    if (!(dist2 > 0.0f)) continue;
                    L = vnorm(L);
                    const float L3[3] = {L.x, L.y, L.z};
                    float ts = 0.0f, tsb = 0.0f;
                    const int32_t is = ora_trace_ray(o3, L3, sc.s.data(), n, &ts);
                    const bool occ_ref = is >= 0 && ts > 0.0f && ts * ts < dist2;
                    const int32_t isb = trace_bvh(B, sph, o, L, dist2, &tsb, q);
                    const bool occ = isb >= 0 && tsb > 0.0f && tsb * tsb < dist2;
                    if (occ != occ_ref) {
                        std::fprintf(stderr, "FAIL [%s] shadow: o=(%a,%a,%a) L=(%a,%a,%a) d2=%a "
                                     "oracle=%d t=%a bvh=%d t=%a\n", sc.name, o.x, o.y, o.z, L.x, L.y,
                                     L.z, dist2, is, ts, isb, tsb);
                        return 1;
                    }
                    occluded += occ_ref;
                }
        std::printf("%-20s spheres=%4u big=%2zu nodes=%5u rays=%7" PRIu64 " hits=%7" PRIu64
                    " occluded=%7" PRIu64 " tests/query=%.1f boxes/query=%.1f exact=%" PRIu64 "\n",
                    sc.name, n, B.big.size(), B.n_nodes, nrays, hits, occluded,
                    (double)q.tests / (2.0 * nrays), (double)q.boxes / (2.0 * nrays), q.exact);
    }
    return rc;
}
