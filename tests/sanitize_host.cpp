// sanitize_host.cpp -- drives the host-side code that reads untrusted input
// or builds data structures (scene.json reader/writer and flattening in
// ataraxia_amd/csrc/scene_graph.cpp, the sphere tree of bvh_build.cpp, and
// the CPU oracle) under AddressSanitizer and UndefinedBehaviorSanitizer.
// tests/test_sanitizers.py builds it with g++ -fsanitize=address,undefined.
//
//   sanitize_host <scene.json>
//
// 1. the file parses, writes back byte for byte, flattens;
// 2. every proper prefix of it is rejected cleanly (ATX_ERR_IO);
// 3. 3000 seeded random mutations (byte flips, inserted structural
//    characters, deleted spans) parse or fail without memory errors, and
//    those that parse also write, re-parse and flatten;
// 4. pathological inputs: deep nesting, long numbers and strings, huge
//    exponents, a graph with many nodes;
// 5. trees over random, duplicated, degenerate and non-finite spheres;
// 6. the oracle renders frames and per-pixel traces of a small scene.
// Exits non-zero (or the sanitizer aborts) on the first failure.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../ataraxia_amd/csrc/bvh_build.h"
#include "../include/atx.h"
extern "C" {
#include "../oracle/atx_oracle.h"
}

namespace {

uint64_t g_state = 0x9E3779B97F4A7C15ull;
uint64_t next_u64()
{
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double uniform() { return (double)(next_u64() >> 11) * (1.0 / 9007199254740992.0); }

int fail(const char* what)
{
    std::fprintf(stderr, "sanitize_host: %s (%s)\n", what, atx_last_error());
    return 1;
}

// Parses; on success also writes, re-parses and flattens.  Returns the
// status of the first parse.
int exercise(const std::string& text)
{
    atx_scene_graph* g = nullptr;
    const int rc = atx_scene_parse_json(text.data(), text.size(), &g);
    if (rc != ATX_OK) {
        if (g) return fail("failed parse returned a graph");
        return rc;
    }
    uint64_t len = 0;
    if (atx_scene_to_json(g, nullptr, 0, &len) != ATX_OK) { atx_scene_destroy(g); return fail("to_json size"); }
    std::vector<char> buf(len + 1);
    if (atx_scene_to_json(g, buf.data(), buf.size(), &len) != ATX_OK) { atx_scene_destroy(g); return fail("to_json"); }
    atx_scene_graph* g2 = nullptr;
    if (atx_scene_parse_json(buf.data(), len, &g2) != ATX_OK) { atx_scene_destroy(g); return fail("re-parse"); }
    uint64_t n = 0;
    atx_scene_flatten(g2, nullptr, 0, &n);
    std::vector<atx_sphere> sph(n + 1);
    if (atx_scene_flatten(g2, sph.data(), sph.size(), &n) != ATX_OK) {
        atx_scene_destroy(g2);
        atx_scene_destroy(g);
        return fail("flatten");
    }
    uint64_t nodes = 0, spheres = 0, mats = 0, lights = 0;
    atx_scene_counts(g2, &nodes, &spheres, &mats, &lights);
    for (uint32_t k = 0; k < nodes; ++k) {
        char name[16];
        float p[3], q[4], s[3];
        uint32_t parent = 0;
        uint64_t ns = 0;
        atx_scene_get_node(g2, k, name, sizeof(name), p, q, s, &parent, &ns);
    }
    atx_scene_destroy(g2);
    atx_scene_destroy(g);
    return ATX_OK;
}

std::string read_file(const char* path)
{
    std::string s;
    FILE* f = std::fopen(path, "rb");
    if (!f) return s;
    char b[4096];
    size_t k;
    while ((k = std::fread(b, 1, sizeof(b), f)) > 0) s.append(b, k);
    std::fclose(f);
    return s;
}

int scene_json(const char* path)
{
    const std::string text = read_file(path);
    if (text.empty()) return fail("cannot read the sample scene");
    if (exercise(text) != ATX_OK) return fail("sample scene does not parse");
    atx_scene_graph* g = nullptr;
    atx_scene_parse_json(text.data(), text.size(), &g);
    uint64_t len = 0;
    atx_scene_to_json(g, nullptr, 0, &len);
    std::vector<char> buf(len + 1);
    atx_scene_to_json(g, buf.data(), buf.size(), &len);
    atx_scene_destroy(g);
    if (std::string(buf.data(), len) != text) return fail("sample scene does not round-trip");

    for (size_t k = 0; k < text.size(); ++k) {
        const std::string pre = text.substr(0, k);
        // a prefix that ends in whitespace after the closing brace does not exist: every one is malformed
        if (exercise(pre) != ATX_ERR_IO) return fail("a truncated file was not rejected");
    }
    static const char kInsert[] = "{}[]\",:-+.eE0123456789 tfn\\u\n";
    int parsed = 0;
    for (int it = 0; it < 3000; ++it) {
        std::string m = text;
        const int edits = 1 + (int)(next_u64() % 4);
        for (int e = 0; e < edits && !m.empty(); ++e) {
            const size_t at = (size_t)(next_u64() % m.size());
            switch (next_u64() % 4) {
            case 0: m[at] = (char)(next_u64() & 0xFF); break;
            case 1: m.insert(at, 1, kInsert[next_u64() % (sizeof(kInsert) - 1)]); break;
            case 2: m.erase(at, 1 + (size_t)(next_u64() % 16)); break;
            default: m.insert(at, m.substr((size_t)(next_u64() % m.size()), 1 + (size_t)(next_u64() % 64))); break;
            }
        }
        const int rc = exercise(m);
        if (rc == ATX_OK) ++parsed;
        else if (rc != ATX_ERR_IO) return fail("a mutated file gave an unexpected status");
    }
    std::printf("scene.json: %zu prefixes rejected, 3000 mutations (%d parsed)\n", text.size(), parsed);

    // pathological documents
    std::string deep(100000, '[');
    if (exercise(deep) != ATX_ERR_IO) return fail("deep nesting");
    if (exercise("{\"camera\": " + std::string(5000, '9') + "}") != ATX_ERR_IO) return fail("long number");
    if (exercise("[1e999999999, -1e-999999999, 0.000000000000000000000000000000001]") != ATX_ERR_IO)
        return fail("extreme exponents in a non-scene document");
    if (exercise("\"" + std::string(100000, 'a') + "\\u12") != ATX_ERR_IO) return fail("unterminated escape");
    std::string many = "{\"camera\":{\"direction\":[0,0,-1],\"fov\":45,\"position\":[0,0,5]},\"lights\":[],"
                       "\"materials\":[],\"sceneGraph\":{\"name\":\"r\",\"transformation\":{\"position\":[0,0,0],"
                       "\"rotation\":[0,0,0,1],\"scale\":[1,1,1]},\"children\":[";
    for (int k = 0; k < 2000; ++k)
        many += std::string(k ? "," : "") + "{\"name\":\"n\",\"transformation\":{\"position\":[1,2,3],\"rotation\":"
                "[0,0,0,1],\"scale\":[1,1,1]},\"spheres\":[{\"center\":[0,0,0],\"radius\":1,\"materialIndex\":7}]}";
    many += "]},\"settings\":{\"accumulation\":true,\"maxBounces\":3,\"skyLight\":false}}";
    if (exercise(many) != ATX_OK) return fail("a 2000-node graph");
    return 0;
}

int trees()
{
    for (int trial = 0; trial < 40; ++trial) {
        const uint32_t n = (uint32_t)(next_u64() % 3000);
        std::vector<float> sph(4 * (size_t)n + 4);
        for (uint32_t i = 0; i < n; ++i) {
            float* s = &sph[4 * i];
            const double kind = uniform();
            s[0] = (float)(uniform() * 40 - 20);
            s[1] = (float)(uniform() * 6);
            s[2] = (float)(uniform() * 40 - 20);
            const float r = (float)(0.3 + uniform() * 1.2);
            s[3] = r * r;
            if (kind < 0.05 && i > 0) std::memcpy(s, &sph[4 * (i - 1)], 16);     // duplicate
            else if (kind < 0.07) s[3] = 0.0f;                                  // degenerate
            else if (kind < 0.08) s[3] = 1e6f;                                  // large (ground-like)
            else if (kind < 0.085) s[0] = NAN;                                   // non-finite
        }
        for (uint32_t layouts : {1u, 8u}) {
            atx::HostBvh t;
            atx::build_bvh(sph.data(), n, layouts, &t);
            if (t.n_nodes && t.nodes.size() != (size_t)t.layouts * t.n_nodes * 4) return fail("tree layout size");
        }
    }
    std::printf("trees: 40 scenes x 2 layouts built\n");
    return 0;
}

int oracle()
{
    ora_sphere s[5] = {{{0, -1000, 0}, 1000, 0}, {{0, 1, 0}, 1, 1}, {{2, 0.5f, 1}, 0.5f, 2},
                       {{-2, 0.7f, -1}, 0.7f, 1}, {{0, 1, 0}, 1, 2}};
    ora_material m[3] = {};
    for (int k = 0; k < 3; ++k) {
        m[k].albedo[0] = m[k].albedo[1] = m[k].albedo[2] = 0.3f + 0.3f * (float)k;
        m[k].roughness = 0.2f + 0.3f * (float)k;
        m[k].metallic = k == 1 ? 1.0f : 0.0f;
        m[k].F0[0] = m[k].F0[1] = m[k].F0[2] = 0.04f;
    }
    m[2].emission_color[0] = 1.0f;
    m[2].emission_intensity = 2.0f;
    ora_light l[2] = {{{5, 10, 5}, {1, 1, 1}, 10}, {{-5, 8, 2}, {1, 0.5f, 0.5f}, 6}};
    ora_scene sc = {s, 5, m, 3, l, 2};
    ora_settings st = {1, 1, {0, 0}, 6};
    ora_camera cam = {{0, 2, 8}, {0, -0.2f, -1}, 45, 0.1f, 100, 33, 17};
    const size_t n = 33 * 17;
    std::vector<float> acc(4 * n), comp(4 * n), rad(4 * n);
    std::vector<uint32_t> img(n);
    uint64_t rays = 0;
    for (uint32_t f = 1; f <= 3; ++f)
        ora_render_frame(&cam, &sc, &st, f, acc.data(), comp.data(), img.data(), rad.data(), 0, 17, &rays);
    if (ora_render_frames_mt(&cam, &sc, &st, 1, 2, acc.data(), nullptr, img.data(), 3, &rays) != 0)
        return fail("oracle threads");
    std::vector<uint32_t> px(n);
    for (size_t k = 0; k < n; ++k) px[k] = (uint32_t)k;
    std::vector<float> pacc(4 * n);
    if (ora_accumulate_pixels(&cam, &sc, &st, 1, 3, px.data(), n, 1, pacc.data(), img.data(), 4) != 0)
        return fail("oracle pixel threads");
    std::vector<int32_t> hit(6 * n), shadow(6 * n);
    std::vector<uint32_t> prays(n);
    ora_render_frame_debug(&cam, &sc, &st, 2, rad.data(), hit.data(), shadow.data(), 6, prays.data(), 0, 17);
    int32_t h[6], sh[6], li[6];
    uint32_t seeds[6];
    float rr[12], dirs[18], out[4];
    ora_path_debug dbg = {h, sh, 0, 0, seeds, li, rr, dirs};
    float d[3] = {0.1f, -0.2f, -1.0f};
    ora_per_pixel(3, 4, 33, d, cam.position, &sc, 5, &st, out, &dbg);
    std::printf("oracle: frames, threads, pixel lists and traces ran (%llu rays)\n", (unsigned long long)rays);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <scene.json>\n", argv[0]);
        return 2;
    }
    if (scene_json(argv[1])) return 1;
    if (trees()) return 1;
    if (oracle()) return 1;
    std::printf("sanitize_host: ok\n");
    return 0;
}
