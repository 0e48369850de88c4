// autotune.cpp -- picks the batch renderer's chunk size and persistent grid by
// timing candidates on the batch itself, and caches the choice.
//
// Candidates: chunk_px in {64, 128, 256} x frames per launch in {16, 32} x
// grid in {full, 3/4} of the device's resident-workgroup capacity.  Each
// candidate renders the first F frames of a spread-out subsample of at most 8
// of the batch's scenes twice (the second run is timed with HIP events; the
// graph is re-captured per candidate) and the one with the highest samples/ms
// wins.  The cache is a text file of
// "<key>\t<chunk_px>\t<workgroups>\t<frames_per_launch>" lines; the key
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

static bool cur_kahan(atx_batch* h);

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

namespace atx {
static bool cur_kahan(atx_batch* h) { return batch_is_kahan(batch_cast(h)); }
}  // namespace atx

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
            unsigned chunk = 0, wg = 0, fpl = 0;
            if (std::getline(ls, k, '\t') && (ls >> chunk >> wg >> fpl) && k == key && chunk % 64 == 0 && chunk) {
                atx_batch_tuning t{chunk, wg, 1, fpl};
                rc = atx_batch_set_tuning(h, &t);
                if (rc != ATX_OK) return rc;
                if (chosen) *chosen = t;
                return ATX_OK;
            }
        }
    }

    // candidates: chunk size x frames per launch x grid fill, each timed on a
    // spread-out subsample of at most 8 scenes (their first F frames, run twice,
    // the second run timed)
    std::vector<uint32_t> sample;
    const size_t total = batch_size(b);
    const size_t step = total > 8 ? total / 8 : 1;
    for (size_t i = 0; i < total && sample.size() < 8; i += step) sample.push_back((uint32_t)i);
    batch_set_subset(b, sample);
    const uint32_t chunks[] = {64u, 128u, 256u};
    const uint32_t frames[] = {16u, 32u};
    double best_rate = -1.0;
    atx_batch_tuning best{kDefaultChunkPx, 0, 1, kDefaultFramesPerLaunch};
    for (uint32_t c : chunks) {
        const uint32_t full = wavefront_full_grid(c, cur_kahan(h));
        const uint32_t grids[] = {full, (full * 3) / 4};
        for (uint32_t F : frames) {
            for (uint32_t g : grids) {
                if (g == 0) continue;
                atx_batch_tuning t{c, g, 1, F};
                rc = atx_batch_set_tuning(h, &t);
                if (rc != ATX_OK) return rc;
                double ms = 0.0;
                rc = batch_run(b, F, &ms);   // warm-up (also captures the graph)
                if (rc == ATX_OK) rc = batch_run(b, F, &ms);
                if (rc != ATX_OK) {
                    batch_set_subset(b, {});
                    return rc;
                }
                atx_render_stats st{};
                atx_batch_get_stats(h, &st);
                const double rate = ms > 0.0 ? (double)st.samples / ms : 0.0;   // samples per ms
                if (rate > best_rate) {
                    best_rate = rate;
                    best = t;
                }
            }
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
        out << key << '\t' << best.chunk_px << '\t' << best.workgroups << '\t' << best.frames_per_launch << '\n';
    }
    return ATX_OK;
}
