// Host-side checks of the persistent kernel's launch layout
// (ataraxia_amd/csrc/wavefront_host.h): chunk ids, pixel and sample offsets,
// launch totals, the 32-bit guard, the flat-loop test of a table, the job
// order (job_pixel_frame), the chunk size chosen per launch (auto_chunk_px),
// the first-bounce cache's layout and hit lists, and the hand-out rules of
// the device-block pool (host_common.h: DevPool).
// Built host-only (hipcc --offload-host-only) by tests/test_wavefront_host.py.
#include <cstdio>
#include <vector>

#include "wavefront_host.h"

static int failures = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::fprintf(stderr, "FAIL line %d: %s\n", __LINE__, #c);     \
            ++failures;                                                   \
        }                                                                 \
    } while (0)

static atx::WfScene scene(uint32_t w, uint32_t h, uint32_t frames, uint32_t nodes)
{
    atx::WfScene s{};
    s.width = w;
    s.height = h;
    s.frames = frames;
    s.bvh.n_nodes = nodes;
    return s;
}

int main()
{
    using namespace atx;
    float3* base = reinterpret_cast<float3*>(uintptr_t(1) << 20);

    // two scenes: 100x50 px x 3 frames (15000 jobs), 7x3 px x 64 frames (1344 jobs)
    std::vector<WfScene> t = {scene(100, 50, 3, 9), scene(7, 3, 64, 0)};
    WfPass p{};
    CHECK(layout_launch(t.data(), t.size(), 128, base, &p));
    CHECK(t[0].chunk_begin == 0 && t[0].n_chunks == 118);      // ceil(15000 / 128)
    CHECK(t[1].chunk_begin == 118 && t[1].n_chunks == 11);     // ceil(1344 / 128)
    CHECK(p.total_chunks == 129 && p.chunk_px == 128);
    CHECK(t[0].px_begin == 0 && t[1].px_begin == 5000 && p.total_px == 5021);
    CHECK(t[0].samples == base && t[1].samples == base + 15000);
    CHECK(launch_samples(t.data(), t.size()) == 15000 + 1344);
    CHECK(!wf_flat_only(t.data(), t.size()) && wf_flat_only(t.data() + 1, 1));

    // every job of a scene belongs to exactly one chunk: jobs [c*P, min((c+1)*P, jobs))
    for (const WfScene& s : t) {
        const uint64_t jobs = (uint64_t)s.width * s.height * s.frames;
        CHECK((uint64_t)s.n_chunks * 128 >= jobs && (uint64_t)(s.n_chunks - 1) * 128 < jobs);
    }

    // a scene of 2^31 or more jobs is refused (the kernel's job indices are 32-bit)
    std::vector<WfScene> big = {scene(32768, 16384, 4, 1)};   // 2^31 jobs
    CHECK(!layout_launch(big.data(), big.size(), 512, base, &p));
    big[0].frames = 3;                                          // 1.6e9 jobs: fits
    CHECK(layout_launch(big.data(), big.size(), 512, base, &p));
    CHECK(big[0].n_chunks == (uint32_t)((3ull * 32768 * 16384 + 511) / 512));

    // the launch's pixel total is guarded too
    std::vector<WfScene> many(3, scene(32768, 32767, 1, 0));   // 3 x ~1.07e9 pixels
    CHECK(!layout_launch(many.data(), many.size(), 512, base, &p));

    // a single 1x1 scene of one frame: one chunk
    std::vector<WfScene> one = {scene(1, 1, 1, 0)};
    CHECK(layout_launch(one.data(), one.size(), 64, base, &p));
    CHECK(one[0].n_chunks == 1 && p.total_chunks == 1 && p.total_px == 1);

    // job order: every (pixel, frame) pair exactly once, for group-size
    // multiples and remainders; 8 pixels of one frame side by side
    const uint32_t sizes[][2] = {{1, 1}, {5, 3}, {8, 64}, {13, 7}, {1000, 1}, {1923, 64}, {4097, 5}};
    for (const auto& sz : sizes) {
        const uint32_t npx = sz[0], frames = sz[1];
        std::vector<uint8_t> seen((size_t)npx * frames, 0);
        bool ok = true;
        for (uint32_t j = 0; j < npx * frames; ++j) {
            uint32_t px = 0, f = 0;
            job_pixel_frame(j, npx, frames, px, f);
            if (px >= npx || f >= frames || seen[(size_t)f * npx + px]++) { ok = false; break; }
            // job_of inverts it (samples are stored by job)
            if (job_of(px, f, npx, frames) != j) { ok = false; break; }
        }
        CHECK(ok);
    }
    // a pixel's frames lie min(8, pixels left in its group) jobs apart, which
    // is the stride k_accumulate reads with
    for (const auto& sz : sizes) {
        const uint32_t npx = sz[0], frames = sz[1];
        bool ok = true;
        for (uint32_t k = 0; k < npx && ok; ++k) {
            const uint32_t left = npx - (k & ~7u), step = left < 8 ? left : 8;
            for (uint32_t f = 0; f < frames; ++f)
                if (job_of(k, f, npx, frames) != job_of(k, 0, npx, frames) + f * step) { ok = false; break; }
        }
        CHECK(ok);
    }
    // division by invariant divisors (multiply and shifts) against / and %
    {
        std::vector<uint32_t> ds;
        for (uint32_t d = 1; d <= 300; ++d) ds.push_back(d);
        for (uint32_t f = 1; f <= 4096; f = f * 3 / 2 + 1) ds.push_back(8 * f);
        const uint32_t big[] = {1023u, 1024u, 1025u, 65535u, 65536u, 65537u, 0x7fffffffu, 0x80000000u,
                                0x80000001u, 0xaaaaaaabu, 0xfffffffeu, 0xffffffffu};
        for (uint32_t d : big) ds.push_back(d);
        uint64_t x = 0x9e3779b97f4a7c15ull;
        for (int k = 0; k < 200; ++k) {
            x = x * 6364136223846793005ull + 1442695040888963407ull;
            ds.push_back((uint32_t)(x >> 33) | 1u);
            ds.push_back((uint32_t)(x >> (40 + k % 20)) + 1u);
        }
        bool ok = true;
        for (uint32_t d : ds) {
            const UDiv v = udiv_make(d);
            std::vector<uint32_t> ns = {0u, 1u, d - 1u, d, d + 1u, 2u * d - 1u, 2u * d, 0x7fffffffu,
                                        0x80000000u, 0xfffffffeu, 0xffffffffu, 0xffffffffu - d};
            const uint64_t q = 0xffffffffull / d;
            ns.push_back((uint32_t)(q * d));
            ns.push_back((uint32_t)(q * d - 1u));
            for (int k = 0; k < 2000; ++k) {
                x = x * 6364136223846793005ull + 1442695040888963407ull;
                ns.push_back((uint32_t)(x >> 32));
                ns.push_back((uint32_t)(x >> (44 + k % 16)));
            }
            for (uint32_t n : ns)
                if (udiv(n, v) != n / d || umod(n, v) != n % d) ok = false;
        }
        CHECK(ok);
        // the job mapping through the divider equals the plain one
        for (const auto& sz : sizes) {
            const uint32_t npx = sz[0], frames = sz[1];
            const UDiv g = udiv_make(kJobGroupPx * frames);
            for (uint32_t j = 0; j < npx * frames && ok; ++j) {
                uint32_t p0, f0, p1, f1;
                job_pixel_frame(j, npx, frames, p0, f0);
                job_pixel_frame(j, npx, g, p1, f1);
                ok = p0 == p1 && f0 == f1;
            }
        }
        CHECK(ok);
    }
    // a block of 64 consecutive jobs is 768 contiguous bytes of samples
    CHECK(64 * kSampleBytes % 128 == 0 && kMaxChunkPx % 64 == 0);
    for (uint32_t j = 0; j < 64; ++j) {
        uint32_t px = 0, f = 0;
        job_pixel_frame(8 * 64 * 3 + j, 1920 * 1080, 64, px, f);
        CHECK(px == 24 + j % 8 && f == j / 8);
    }
    {
        uint32_t px = 0, f = 0;
        job_pixel_frame(13 * 7 - 1, 13, 7, px, f);   // last group: pixels 8..12, 5 wide
        CHECK(px == 12 && f == 6);
        job_pixel_frame(8 * 7 + 5, 13, 7, px, f);
        CHECK(px == 8 && f == 1);
    }

    // chunk size chosen per launch: 1080p frames on a 256-CU grid
    // (tree launches 6 workgroups per CU, flat-loop launches 7)
    const uint64_t f1080 = 1920ull * 1080;
    CHECK(auto_chunk_px(f1080, 1536) == 256);          // 337 jobs per wave
    CHECK(auto_chunk_px(8 * f1080, 1536) == 256);      // 2700
    CHECK(auto_chunk_px(16 * f1080, 1536) == 512);     // 5400
    CHECK(auto_chunk_px(64 * f1080, 1536) == 512);
    CHECK(auto_chunk_px(16 * f1080, 1792) == 256);     // 4628
    CHECK(auto_chunk_px(32 * f1080, 1792) == 512);
    CHECK(auto_chunk_px(1, 0) == 256 && auto_chunk_px(0, 1) == 256);
    for (uint64_t jobs : {uint64_t(1), uint64_t(1000), f1080, 64 * f1080, uint64_t(1) << 31}) {
        const uint32_t c = auto_chunk_px(jobs, 1536);
        CHECK(c % 64 == 0 && c >= 64 && c <= kMaxChunkPx);
    }

    // first-bounce cache layout: records, light terms, hit list and its
    // length lie one after another, 16-byte aligned, inside first_cache_bytes
    for (size_t npx : {size_t(1), size_t(7), size_t(1920) * 1080, size_t(4097)})
        for (uint32_t lights : {0u, 1u, 8u, 32u, 33u}) {
            char* b0 = reinterpret_cast<char*>(uintptr_t(1) << 24);
            const FirstCacheView v = first_cache_view(b0, npx, lights);
            const char* rec = reinterpret_cast<const char*>(v.rec);
            const char* lt = reinterpret_cast<const char*>(v.light);
            const char* hits = reinterpret_cast<const char*>(v.hits);
            const char* nh = reinterpret_cast<const char*>(v.n_hits);
            CHECK(rec == b0);
            CHECK((lights >= 1 && lights <= 32) == (lt != nullptr));
            if (lt) CHECK(lt == rec + npx * 64 && hits == lt + npx * lights * 16);
            else CHECK(hits == rec + npx * 64);
            CHECK(nh >= hits + npx * 4 && (uintptr_t)nh % 16 == 0 && (uintptr_t)hits % 16 == 0);
            CHECK(nh + 4 <= b0 + first_cache_bytes(npx, lights));
        }

    // hit lists: only with a cache, at least one bounce and no debug capture;
    // the skipped primary rays are the missing pixels' in every frame
    {
        char* buf = reinterpret_cast<char*>(uintptr_t(1) << 24);
        WfScene w = scene(100, 50, 16, 3);
        w.n_light = 8;
        w.max_bounces = 4;
        use_hit_list(w, buf, 1234);
        CHECK(w.hit_px == nullptr);                       // no cache
        set_first_cache(w, buf);
        use_hit_list(w, buf, 1234);
        CHECK(w.hit_px == first_cache_view(buf, 5000, 8).hits && w.n_hit == 1234);
        CHECK(scene_jobs(w) == 1234ull * 16 && skipped_primaries(&w, 1) == (5000ull - 1234) * 16);
        WfScene z = w;
        z.hit_px = nullptr;
        z.max_bounces = 0;
        use_hit_list(z, buf, 1234);
        CHECK(z.hit_px == nullptr && scene_jobs(z) == 5000ull * 16 && skipped_primaries(&z, 1) == 0);
        WfScene d = w;
        d.hit_px = nullptr;
        d.dbg_hit = reinterpret_cast<int32_t*>(buf);
        use_hit_list(d, buf, 1234);
        CHECK(d.hit_px == nullptr);
        std::vector<WfScene> two = {w, z};
        CHECK(launch_jobs(two.data(), 2) == 1234ull * 16 + 5000ull * 16);
        CHECK(layout_launch(two.data(), 2, 512, base, &p));
        CHECK(two[0].n_chunks == (1234u * 16 + 511) / 512 && two[1].chunk_begin == two[0].n_chunks);
        CHECK(two[1].samples == base + 5000 * 16);   // samples stay per pixel
    }

    // device-block pool: a block goes to a request of at least half its size
    // (+64 KiB) whose excess fits the slack.  Fake addresses; every block is
    // taken back before the pool is destroyed, so nothing is freed.
    {
        DevPool pool(size_t(1) << 40);
        char* b = reinterpret_cast<char*>(uintptr_t(1) << 30);
        pool.give(b, 1000);
        pool.give(b + 4096, 300000);
        size_t cap = 0;
        CHECK(pool.pooled() == 301000);
        CHECK(pool.take(400000, &cap) == nullptr);                      // no block that large
        CHECK(pool.take(100000, &cap) == nullptr);                      // 300000 > 2 * 100000 + 64 KiB
        pool.set_slack(100000);
        CHECK(pool.take(190000, &cap) == nullptr);                      // excess 110000 > slack
        CHECK(pool.take(200000, &cap) == b + 4096 && cap == 300000);    // excess 100000 = slack
        CHECK(pool.take(900, &cap) == nullptr);                         // slack used up
        pool.set_slack(SIZE_MAX);
        CHECK(pool.take(900, &cap) == b && cap == 1000);
        CHECK(pool.pooled() == 0);
    }

    if (failures == 0) std::printf("wavefront layout: all checks passed\n");
    return failures == 0 ? 0 : 1;
}
