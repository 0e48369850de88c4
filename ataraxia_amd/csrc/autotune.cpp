// autotune.cpp -- picks the batch renderer's chunk size, frames per launch,
// launch bounds (the path kernel's register budget: resident waves per
// SIMD) and persistent grid by timing candidates on the batch itself, and
// caches the choice.
//
// Candidates: chunk_px in {256, 512} x frames per launch in {32, 64, 128, 256}
// with the default kernel instance on its full resident-workgroup grid; then
// the winner with each other instance (5 to 8 waves per SIMD) on that
// instance's full grid; then the winner on 3/4 of its grid.
// Each candidate renders the first F frames of a spread-out subsample of the
// batch's scenes (up to 16 Mpixel, at least 8 scenes) twice (the second run is
// timed with HIP events; the graph is re-captured per candidate) and the one
// with the highest samples/ms wins.  The cache is a text file of
// "<key>\t<chunk_px>\t<workgroups>\t<frames_per_launch>\t<waves_per_simd>" lines; the key
// names the device (arch, CU count) and the workload (scene count, pixels,
// spheres, spp, bounces).
#include <stdio.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "batch_internal.h"
#include "host_common.h"
#include "wavefront_host.h"

namespace atx {

static std::string device_key()
{
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return "unknown";
    std::ostringstream os;
    os << prop.gcnArchName << "/cu" << prop.multiProcessorCount;
    std::string k = os.str();
    for (char& c : k)
        if (c == '\t' || c == '\n' || c == ' ') c = '_';
    return k;
}

}  // namespace atx

using namespace atx;

extern "C" int atx_batch_autotune(atx_batch* h, const char* cache_path, atx_batch_tuning* chosen)
{
    Batch* b = batch_cast(h);
    ATX_CHECK_ARG(b, "batch is NULL");
    atx_batch_tuning cur{};
    int rc = atx_batch_get_tuning(h, &cur);
    if (rc != ATX_OK) return rc;

    // workload signature from the public per-scene state
    uint64_t pixels = 0, spheres = 0, spp = 0, bounces = 0, n = 0;
    batch_signature(b, &n, &pixels, &spheres, &spp, &bounces);
    std::ostringstream ks;
    ks << device_key() << "|n" << n << "|px" << pixels << "|sph" << spheres << "|spp" << spp << "|b" << bounces << "|acc"
       << batch_accel(b) << "|k" << (batch_is_kahan(b) ? 1 : 0);
    const std::string key = ks.str();

    if (cache_path) {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string k;
            unsigned chunk = 0, wg = 0, fpl = 0, waves = 0;
            if (std::getline(ls, k, '\t') && (ls >> chunk >> wg >> fpl >> waves) && k == key && chunk % 64 == 0 &&
                chunk) {
                atx_batch_tuning t{chunk, wg, 1, fpl, waves};
                rc = atx_batch_set_tuning(h, &t);
                if (rc != ATX_OK) return rc;
                if (chosen) *chosen = t;
                return ATX_OK;
            }
        }
    }

    // Candidates: chunk size x frames per launch on the full grid, then the
    // winner on 3/4 of it.  Each is timed on a spread-out subsample of the
    // scenes (every k-th scene until kTunePixels pixels, at least 8 scenes):
    // their first F frames, run twice, the second run timed.  A launch's tail
    // and scene switches weigh in the timing as they do in the full batch.
    constexpr uint64_t kTunePixels = 16ull << 20;
    std::vector<uint32_t> sample;
    const size_t total = batch_size(b);
    for (size_t stride = 1; stride <= total; stride *= 2) {
        std::vector<uint32_t> pick;
        uint64_t px = 0;
        for (size_t i = 0; i < total; i += stride) {
            pick.push_back((uint32_t)i);
            px += batch_scene_pixels(b, (uint32_t)i);
        }
        sample = pick;
        if (px <= kTunePixels || pick.size() <= 8) break;
    }
    batch_set_subset(b, sample);
    double best_rate = -1.0;
    atx_batch_tuning best{0, 0, 1, kDefaultFramesPerLaunch, 0};
    auto try_one = [&](const atx_batch_tuning& t) -> int {
        atx_batch_tuning tt = t;
        int e = atx_batch_set_tuning(h, &tt);
        if (e != ATX_OK) return e;
        double ms = 0.0;
        e = batch_run(b, t.frames_per_launch, &ms);   // warm-up (also captures the graph)
        if (e == ATX_OK) e = batch_run(b, t.frames_per_launch, &ms);
        if (e != ATX_OK) return e;
        atx_render_stats st{};
        atx_batch_get_stats(h, &st);
        const double rate = ms > 0.0 ? (double)st.samples / ms : 0.0;   // samples per ms
        if (rate > best_rate) {
            best_rate = rate;
            best = t;
        }
        return ATX_OK;
    };
    const bool flat = batch_flat_only(b);
    const uint32_t w0 = wf_waves(0, flat);
    const uint32_t chunks[] = {256u, 512u};
    const uint32_t frames[] = {32u, 64u, 128u, 256u};
    for (uint32_t c : chunks) {
        for (uint32_t F : frames) {
            rc = try_one(atx_batch_tuning{c, wavefront_full_grid(flat, w0), 1, F, w0});
            if (rc != ATX_OK) {
                batch_set_subset(b, {});
                return rc;
            }
        }
    }
    for (uint32_t w = kWfWavesMin; w <= kWfWavesMax; ++w) {
        if (w == w0) continue;
        rc = try_one(atx_batch_tuning{best.chunk_px, wavefront_full_grid(flat, w), 1, best.frames_per_launch, w});
        if (rc != ATX_OK) {
            batch_set_subset(b, {});
            return rc;
        }
    }
    const uint32_t three_quarters = (best.workgroups * 3) / 4;
    if (three_quarters > 0) {
        rc = try_one(atx_batch_tuning{best.chunk_px, three_quarters, 1, best.frames_per_launch, best.waves_per_simd});
        if (rc != ATX_OK) {
            batch_set_subset(b, {});
            return rc;
        }
    }
    batch_set_subset(b, {});
    rc = atx_batch_set_tuning(h, &best);
    if (rc != ATX_OK) return rc;
    if (chosen) *chosen = best;
    if (cache_path) {
        std::ofstream out(cache_path, std::ios::app);
        if (!out) {
            set_error("cannot write autotune cache %s", cache_path);
            return ATX_ERR_IO;
        }
        out << key << '\t' << best.chunk_px << '\t' << best.workgroups << '\t' << best.frames_per_launch << '\t'
            << best.waves_per_simd << '\n';
    }
    return ATX_OK;
}
