// server.cpp -- continuous-batching render service on one GPU.
//
// Submitters copy their scene into a request and return at once.  One worker
// thread drives the device: it admits queued requests while
// the admitted ones fill less than tick_mpaths million paths per launch (so
// the device stays full without every queued request sharing it) and their
// device buffers fit the HBM budget (geometry, accumulation, optional Kahan
// compensation, RGBA8 image, primary directions, first-bounce cache, the
// tick's samples).  The
// next request is the one with the highest response ratio, (waiting time +
// estimated render time) / estimated render time, with the render time
// estimated from the throughput measured so far: short requests go first and
// a long one's priority grows while it waits.  Every admitted request is
// rendered together with the others by one launch of the persistent
// wavefront kernel per tick (frames_per_tick frames per request, each request
// at its own frame index).  Requests that arrive while a launch runs join at
// the next launch: the worker enqueues a tick and admits the next tick's
// requests while it runs (their primary rays and first-bounce caches are
// computed on a second stream, beside the tick).  Requests that reach their
// spp go to a retire thread, which copies their results to host memory and
// marks them done; the worker then returns their device blocks to a pool
// that later admissions take from (a hipFree would wait for the device),
// which makes room for the next admissions.
// Per-request timestamps (submit, admit, done) give the latency
// distribution.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "device_types.h"
#include "host_common.h"
#include "wavefront_host.h"

namespace atx {

struct Request {
    uint64_t id = 0;
    // host copy of the scene
    std::vector<atx_sphere> spheres;
    std::vector<atx_material> materials;
    std::vector<atx_light> lights;
    atx_settings settings{};
    atx_camera cam{};
    uint32_t spp = 1;
    size_t device_bytes = 0;
    // device state while admitted
    DevScene geo;
    DevBuf accum, comp, image;
    DevBuf dirs;                    // primary ray direction per pixel
    DevBuf first;                   // first-bounce cache per pixel (launch_first_hits)
    uint32_t n_hit = 0;             // pixels of its hit list
    uint32_t frames_done = 0;
    // results
    std::vector<uint32_t> rgba;
    std::vector<float> acc_host;
    int state = ATX_REQ_QUEUED;
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::string error;
};

struct Server {
    // Device blocks of released requests, reused by later ones (first member:
    // destroyed after every request that may still hold a block).
    DevPool pool;
    atx_server_config cfg{};
    std::chrono::steady_clock::time_point epoch;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;          // waiting for admission
    std::vector<std::shared_ptr<Request>> active;        // worker-owned
    std::map<uint64_t, std::shared_ptr<Request>> all;    // every live request by id
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    // finished requests whose results the retire thread copies out
    std::thread retirer;
    std::deque<std::shared_ptr<Request>> retire_q;
    std::vector<std::shared_ptr<Request>> release_q;   // retired, device buffers still held
    std::condition_variable cv_retire;
    bool retire_stop = false;
    atx_server_stats stats{};
    // worker-only device resources
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_prev = nullptr;   // end of the previous launch (gap_ms)
    // admission kernels (primary rays, first-bounce caches) run on their own
    // stream, beside the running launch; the worker waits for them (not for
    // the launch) to read the caches' hit counts
    hipStream_t adm_stream = nullptr;
    bool no_hit_lists = false;      // a hit count could not be read: jobs cover every pixel
    bool have_prev = false;         // the previous launch ran without an idle period since
    DevBuf table, queue_ctr, counters;
    DevBuf samples;                 // the tick's sample buffer (shared by the active requests)
    size_t samples_excess = 0;      // its capacity beyond the active requests' shares, charged to hbm_in_use
    double rendered = 0.0, busy_s = 0.0;   // samples rendered and GPU seconds so far (worker)
    // persistent grids: ticks with trees (6- and 8-wave instances), flat-loop-only ticks
    uint32_t grid = 0, grid8 = 0, grid_flat = 0;
    int device = 0;
};

static double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

static double now_s(const Server* s)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->epoch).count();
}

// Device bytes a request holds while admitted: scene, tree, accumulation,
// compensation, image, primary ray directions, first-bounce cache, and its
// share of the tick's sample buffer.
static size_t request_bytes(const Request& r, bool kahan, uint32_t frames_per_tick)
{
    const size_t npx = (size_t)r.cam.width * r.cam.height;
    const size_t f = std::min<size_t>(frames_per_tick, r.spp);
    return scene_device_bytes(r.spheres.size(), r.materials.size(), r.lights.size()) +
           npx * (16 + (kahan ? 16 : 0) + 4 + 16) + npx * f * kSampleBytes +
           (r.spp >= 2 ? first_cache_bytes(npx, (uint32_t)r.lights.size()) : 0);
}

// Device bytes an admitted request actually holds: its blocks' capacities
// (a pooled block may be up to twice the size asked for) and its share of
// the tick's sample buffer.  Admission checks request_bytes against the
// budget, then charges this.
static size_t held_bytes(const Request& r, uint32_t frames_per_tick)
{
    const size_t npx = (size_t)r.cam.width * r.cam.height;
    const size_t f = std::min<size_t>(frames_per_tick, r.spp);
    return r.geo.bytes() + r.accum.bytes() + r.comp.bytes() + r.image.bytes() + r.dirs.bytes() + r.first.bytes() +
           npx * f * kSampleBytes;
}

// A launch's paths must stay below 2^31 (the kernel's 32-bit job indices).
constexpr uint64_t kMaxTickPaths = 1ull << 31;

// Paths a request adds to the next launch.
static uint64_t tick_paths_of(const Request& r, uint32_t frames_per_tick)
{
    return (uint64_t)r.cam.width * r.cam.height * std::min(frames_per_tick, r.spp - r.frames_done);
}

// Device allocations of one request (worker thread).
static bool admit(Server* s, Request& r)
{
    r.geo.use_pool(&s->pool);
    for (DevBuf* b : {&r.accum, &r.comp, &r.image, &r.dirs, &r.first})
        b->use_pool(&s->pool);
    int rc = upload_scene(r.spheres.data(), r.spheres.size(), r.materials.data(), r.materials.size(),
                          r.lights.data(), r.lights.size(), &r.geo);
    const size_t npx = (size_t)r.cam.width * r.cam.height;
    if (rc == ATX_OK && (r.accum.reserve(npx * 16) != hipSuccess || r.image.reserve(npx * 4) != hipSuccess ||
                         r.dirs.reserve(npx * 16) != hipSuccess ||
                         (r.spp >= 2 && r.first.reserve(first_cache_bytes(npx, (uint32_t)r.lights.size())) != hipSuccess) ||
                         (s->cfg.kahan && r.comp.reserve(npx * 16) != hipSuccess)))
        rc = ATX_ERR_NOMEM;
    if (rc == ATX_OK) {
        // on the admission stream, beside the running launch (admit_queued
        // waits for it before the request's first launch)
        WfScene w;
        fill_wf_scene(r.geo, r.settings, r.cam, nullptr, nullptr, nullptr, false, 1u, 1u, (int)s->cfg.accel, &w);
        w.dirs = r.dirs.as<float4>();
        if (r.spp >= 2) set_first_cache(w, r.first.get());
        // the first-bounce cache pays from two frames per request on
        if (launch_camera_dirs(w, r.dirs.as<float4>(), s->adm_stream) != hipSuccess ||
            (r.spp >= 2 && launch_first_hits(w, s->adm_stream) != hipSuccess)) {
            set_error("primary rays: launch failed");
            rc = ATX_ERR_HIP;
        }
    }
    if (rc != ATX_OK) {
        r.error = atx_last_error();
        // a kernel of this request may be queued on the admission stream:
        // let it finish before its blocks go back to the pool
        (void)hipStreamSynchronize(s->adm_stream);
        r.geo = DevScene();
        r.accum.release();
        r.comp.release();
        r.image.release();
        r.dirs.release();
        r.first.release();
        return false;
    }
    return true;
}

// Releases a request's device buffers.  A release waits for the device to
// be idle, so only the worker calls it, between launches.
static void release_device(Server* s, Request& r)
{
    r.geo = DevScene();
    r.accum.release();
    r.comp.release();
    r.image.release();
    r.dirs.release();
    r.first.release();
    std::lock_guard<std::mutex> lk(s->mu);
    s->stats.hbm_in_use -= r.device_bytes;
}

// Releases the device buffers of requests the retire thread has finished
// (worker, with the device idle).
static void release_returned(Server* s)
{
    std::vector<std::shared_ptr<Request>> rel;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        rel.swap(s->release_q);
    }
    if (rel.empty())
        return;
    const auto t0 = std::chrono::steady_clock::now();
    for (auto& r : rel) release_device(s, *r);
    const double ms = ms_since(t0);
    std::lock_guard<std::mutex> lk(s->mu);
    s->stats.release_ms += ms;
}

// Copies a finished request's results to host memory (ok) and marks it done
// or failed.  The worker releases the device buffers at once (release_now);
// the retire thread hands the request back to the worker for that.
static void finish(Server* s, const std::shared_ptr<Request>& rp, bool ok, bool release_now)
{
    Request& r = *rp;
    const size_t npx = (size_t)r.cam.width * r.cam.height;
    if (ok) {
        r.rgba.resize(npx);
        r.acc_host.resize(4 * npx);
        if (hipMemcpy(r.rgba.data(), r.image.get(), npx * 4, hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(r.acc_host.data(), r.accum.get(), npx * 16, hipMemcpyDeviceToHost) != hipSuccess) {
            ok = false;
            r.error = "result copy failed";
        }
    }
    if (release_now) release_device(s, r);
    std::lock_guard<std::mutex> lk(s->mu);
    if (!release_now) s->release_q.push_back(rp);
    r.state = ok ? ATX_REQ_DONE : ATX_REQ_FAILED;
    r.t_done = now_s(s);
    if (ok) {
        s->stats.completed++;
        s->stats.samples += (uint64_t)npx * r.spp;
    } else {
        s->stats.failed++;
    }
    s->cv_done.notify_all();
}

// Picks queued requests (highest response ratio first) while the resident
// ones fill less than a launch and the budget allows, and makes them
// resident; they join the next launch.
static void admit_queued(Server* s)
{
    std::vector<std::shared_ptr<Request>> newly;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        const uint32_t F = s->cfg.frames_per_tick;
        const uint64_t target =
            s->cfg.tick_mpaths == UINT32_MAX ? UINT64_MAX : (uint64_t)s->cfg.tick_mpaths * 1000000ull;
        uint64_t tick_paths = 0;
        for (const auto& r : s->active) tick_paths += tick_paths_of(*r, F);
        // samples per second: measured, or a nominal rate before the first tick
        const double rate = s->busy_s > 0.0 ? s->rendered / s->busy_s : 4e9;
        const double now = now_s(s);
        while (!s->queue.empty()) {
            const bool idle = s->active.empty() && newly.empty();
            if (!idle && tick_paths >= target) break;
            if (s->cfg.max_active && s->active.size() + newly.size() >= s->cfg.max_active) break;
            size_t pick = 0;
            double best = -1.0;
            for (size_t k = 0; k < s->queue.size(); ++k) {
                const Request& q = *s->queue[k];
                const double est = std::max(1e-6, (double)q.cam.width * q.cam.height * q.spp / rate);
                const double ratio = (now - q.t_submit + est) / est;
                if (ratio > best) {
                    best = ratio;
                    pick = k;
                }
            }
            auto r = s->queue[pick];
            if (s->stats.hbm_in_use + r->device_bytes > s->cfg.hbm_budget_bytes) break;
            if (tick_paths + tick_paths_of(*r, F) >= kMaxTickPaths) break;   // 32-bit job indices
            s->stats.hbm_in_use += r->device_bytes;
            s->stats.hbm_high_water = std::max(s->stats.hbm_high_water, s->stats.hbm_in_use);
            tick_paths += tick_paths_of(*r, F);
            newly.push_back(r);
            s->queue.erase(s->queue.begin() + (ptrdiff_t)pick);
        }
    }
    if (newly.empty()) return;
    const auto t0 = std::chrono::steady_clock::now();
    for (auto& r : newly) {
        {
            // pooled blocks larger than asked for may use only what the
            // budget has left, so the bytes actually held stay within it
            std::lock_guard<std::mutex> lk(s->mu);
            s->pool.set_slack(s->cfg.hbm_budget_bytes > s->stats.hbm_in_use
                                  ? (size_t)(s->cfg.hbm_budget_bytes - s->stats.hbm_in_use) : 0);
        }
        const bool admitted = admit(s, *r);
        s->pool.set_slack(SIZE_MAX);
        if (admitted) {
            std::lock_guard<std::mutex> lk(s->mu);
            const size_t held = held_bytes(*r, s->cfg.frames_per_tick);
            s->stats.hbm_in_use = s->stats.hbm_in_use - r->device_bytes + held;
            s->stats.hbm_high_water = std::max(s->stats.hbm_high_water, s->stats.hbm_in_use);
            r->device_bytes = held;
            r->state = ATX_REQ_RUNNING;
            r->t_admit = now_s(s);
            s->stats.admitted++;
            s->active.push_back(r);
        } else {
            finish(s, r, false, true);
        }
    }
    // the hit counts of the new first-bounce caches (the next launch's layout
    // needs them): wait for the admission stream, not for the running launch
    bool counts_ok = hipStreamSynchronize(s->adm_stream) == hipSuccess;
    for (auto& r : newly) {
        if (r->state != ATX_REQ_RUNNING || r->spp < 2) continue;
        const size_t npx = (size_t)r->cam.width * r->cam.height;
        const FirstCacheView v = first_cache_view(r->first.get(), npx, (uint32_t)r->lights.size());
        counts_ok = counts_ok && hipMemcpy(&r->n_hit, v.n_hits, sizeof(uint32_t), hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (!counts_ok) {
        // without the counts every request renders all of its pixels
        for (auto& r : newly) r->n_hit = 0;
        s->no_hit_lists = true;
    }
    const double ms = ms_since(t0);
    std::lock_guard<std::mutex> lk(s->mu);
    s->stats.admit_ms += ms;
}

// Enqueues one launch over the resident requests (listed in `in_tick`) on the
// worker's stream; does not wait for it.
static bool launch_tick(Server* s, std::vector<WfScene>& table, std::vector<std::shared_ptr<Request>>& in_tick)
{
    const uint32_t F = s->cfg.frames_per_tick;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        in_tick = s->active;
    }
    table.clear();
    for (auto& r : in_tick) {
        WfScene w;
        fill_wf_scene(r->geo, r->settings, r->cam, r->accum.as<float4>(),
                      s->cfg.kahan ? r->comp.as<float4>() : nullptr, r->image.as<uint32_t>(), s->cfg.kahan != 0,
                      r->frames_done + 1u, std::min(F, r->spp - r->frames_done), (int)s->cfg.accel, &w);
        w.dirs = r->dirs.as<float4>();
        if (r->spp >= 2) {
            set_first_cache(w, r->first.get());
            if (!s->no_hit_lists) use_hit_list(w, r->first.get(), r->n_hit);
        }
        table.push_back(w);
    }
    WfPass wp{};
    // The buffer grows to the largest tick since the server last idled (a
    // hipFree and hipMalloc per shrink would stall the device between
    // ticks).  Its capacity beyond the shares the active requests are
    // charged (kSampleBytes per pixel and frame of a tick) is charged to the budget
    // as well, so hbm_in_use counts every byte the server holds.
    const size_t need = std::max<size_t>(1, launch_samples(table.data(), table.size())) * kSampleBytes;
    bool ok = s->samples.reserve(need) == hipSuccess;
    {
        size_t shares = 0;
        for (auto& r : in_tick)
            shares += (size_t)r->cam.width * r->cam.height * std::min<size_t>(F, r->spp) * kSampleBytes;
        const size_t excess = s->samples.bytes() > shares ? s->samples.bytes() - shares : 0;
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.hbm_in_use = s->stats.hbm_in_use - s->samples_excess + excess;
        s->stats.hbm_high_water = std::max(s->stats.hbm_high_water, s->stats.hbm_in_use);
        s->samples_excess = excess;
    }
    const bool flat = wf_flat_only(table.data(), table.size());
    const uint32_t waves = flat ? 0u : wf_tree_waves(launch_jobs(table.data(), table.size()));
    const WfLaunch shape{flat ? s->grid_flat : waves == 8u ? s->grid8 : s->grid, flat, waves};
    const uint32_t P =
        s->cfg.chunk_px ? s->cfg.chunk_px : auto_chunk_px(launch_jobs(table.data(), table.size()), shape.workgroups);
    if (ok) ok = layout_launch(table.data(), table.size(), P, s->samples.as<float3>(), &wp);
    ok = ok && s->table.reserve(table.size() * sizeof(WfScene)) == hipSuccess &&
         hipMemcpyAsync(s->table.get(), table.data(), table.size() * sizeof(WfScene), hipMemcpyHostToDevice,
                        s->stream) == hipSuccess &&
         launch_zero_u32(s->queue_ctr.as<unsigned int>(), 1, s->stream) == hipSuccess &&
         hipEventRecord(s->ev0, s->stream) == hipSuccess;
    if (ok) {
        wp.scenes = s->table.as<WfScene>();
        wp.n_scenes = (uint32_t)table.size();
        wp.queue = s->queue_ctr.as<unsigned int>();
        wp.counters = s->counters.as<unsigned long long>();
        ok = launch_wavefront(wp, shape, s->stream) == hipSuccess &&
             hipEventRecord(s->ev1, s->stream) == hipSuccess;
    }
    return ok;
}

// Waits for the launch, advances the frames of the requests that took part
// and moves the finished ones to `done` (their results are final).
static void end_tick(Server* s, bool ok, const std::vector<WfScene>& table,
                     const std::vector<std::shared_ptr<Request>>& in_tick, std::vector<std::shared_ptr<Request>>& done)
{
    const uint32_t F = s->cfg.frames_per_tick;
    const auto t_wait0 = std::chrono::steady_clock::now();
    ok = hipStreamSynchronize(s->stream) == hipSuccess && ok;
    const double wms = ms_since(t_wait0);
    float ms = 0.0f, gap = 0.0f;
    if (ok) (void)hipEventElapsedTime(&ms, s->ev0, s->ev1);
    if (ok && s->have_prev) (void)hipEventElapsedTime(&gap, s->ev_prev, s->ev0);
    std::swap(s->ev1, s->ev_prev);
    s->have_prev = ok;
    if (ok) {
        s->busy_s += ms * 1e-3;
        for (const WfScene& w : table) s->rendered += (double)w.width * w.height * w.frames;
    }
    size_t free_b = 0, total_b = 0;
    if ((s->stats.ticks & 63u) == 0u) (void)hipMemGetInfo(&free_b, &total_b);
    std::vector<std::shared_ptr<Request>> failed;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.ticks++;
        s->stats.gpu_busy_ms += ms;
        s->stats.wait_ms += wms;
        s->stats.gap_ms += gap;
        if (total_b) s->stats.device_used_high_water = std::max<uint64_t>(s->stats.device_used_high_water, total_b - free_b);
        for (auto& r : in_tick) {
            if (!ok) {
                failed.push_back(r);
                continue;
            }
            r->frames_done += std::min(F, r->spp - r->frames_done);
            if (r->frames_done >= r->spp) done.push_back(r);
        }
        auto gone = [&](const std::shared_ptr<Request>& r) {
            return std::find(done.begin(), done.end(), r) != done.end() ||
                   std::find(failed.begin(), failed.end(), r) != failed.end();
        };
        s->active.erase(std::remove_if(s->active.begin(), s->active.end(), gone), s->active.end());
    }
    for (auto& r : failed) {
        r->error = "launch failed";
        finish(s, r, false, true);
    }
    release_returned(s);   // the device is idle
}

// Hands finished requests to the retire thread, which copies their results
// to host memory and releases their device buffers.
static void retire(Server* s, std::vector<std::shared_ptr<Request>>& done)
{
    if (done.empty()) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (auto& r : done) s->retire_q.push_back(r);
    }
    done.clear();
    s->cv_retire.notify_one();
}

static void retirer_main(Server* s)
{
    (void)hipSetDevice(s->device);
    for (;;) {
        std::shared_ptr<Request> r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_retire.wait(lk, [s] { return s->retire_stop || !s->retire_q.empty(); });
            if (s->retire_q.empty()) break;   // stopping, nothing left
            r = s->retire_q.front();
            s->retire_q.pop_front();
        }
        const auto t0 = std::chrono::steady_clock::now();
        finish(s, r, true, false);
        const double ms = ms_since(t0);
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stats.retire_ms += ms;
        }
        s->cv_work.notify_one();   // its memory can be released and reused
    }
}

// The worker: each pass enqueues a launch over the resident requests and,
// while it runs, admits queued requests for the next launch and hands the
// requests the previous launch finished to the retire thread (neither group
// takes part in the running launch), then waits for the launch.  Uploads,
// result copies and releases therefore overlap rendering instead of idling
// the device.
static void worker_main(Server* s)
{
    (void)hipSetDevice(s->device);
    std::vector<WfScene> table;
    std::vector<std::shared_ptr<Request>> in_tick, done;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop || !s->queue.empty() || !s->active.empty() || !done.empty(); });
            if (s->stop) break;
        }
        bool idle;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            idle = s->active.empty();
        }
        if (idle) {
            release_returned(s);   // nothing runs: free retired requests' memory, then admit
            s->samples.release();  // and the tick's sample buffer, sized for ticks that have ended
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->stats.hbm_in_use -= s->samples_excess;
                s->samples_excess = 0;
            }
            admit_queued(s);
            s->have_prev = false;  // the device idled for want of work: no gap to count
        }
        bool launched = false, ok = false;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            launched = !s->active.empty();
        }
        if (launched) {
            ok = launch_tick(s, table, in_tick);
            if (ok) admit_queued(s);   // joins the next launch
        }
        retire(s, done);
        if (launched) {
            end_tick(s, ok, table, in_tick, done);
        } else {
            // nothing could be admitted: wait until the retire thread hands
            // memory back (or something else changes) instead of spinning
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait_for(lk, std::chrono::milliseconds(5),
                                [s] { return s->stop || !s->release_q.empty() || s->queue.empty(); });
        }
    }
    retire(s, done);
}

static Server* cast(atx_server* h) { return reinterpret_cast<Server*>(h); }

}  // namespace atx

using namespace atx;

extern "C" {

int atx_server_create(atx_server** out, const atx_server_config* cfg)
{
    ATX_CHECK_ARG(out, "out is NULL");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        set_error("no HIP device available");
        return ATX_ERR_NODEVICE;
    }
    std::unique_ptr<Server> s(new (std::nothrow) Server());
    if (!s) { set_error("out of host memory"); return ATX_ERR_NOMEM; }
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.frames_per_tick) s->cfg.frames_per_tick = kDefaultFramesPerLaunch;
    if (!s->cfg.tick_mpaths) s->cfg.tick_mpaths = 128;
    ATX_CHECK_ARG(s->cfg.chunk_px % 64 == 0 && s->cfg.chunk_px <= kMaxChunkPx, "chunk_px must be a multiple of 64, at most 1024");
    ATX_CHECK_ARG(s->cfg.frames_per_tick <= 4096, "frames_per_tick must be <= 4096");
    s->cfg.kahan = s->cfg.kahan ? 1u : 0u;
    ATX_CHECK_ARG(s->cfg.accel == ATX_ACCEL_BVH || s->cfg.accel == ATX_ACCEL_NONE, "unknown accel %u",
                  s->cfg.accel);
    ATX_HIP(hipGetDevice(&s->device));
    if (!s->cfg.hbm_budget_bytes) {
        size_t free_b = 0, total_b = 0;
        ATX_HIP(hipMemGetInfo(&free_b, &total_b));
        s->cfg.hbm_budget_bytes = free_b / 2;
    }
    // released requests' blocks kept for reuse: at most a quarter of the budget
    s->pool.set_limit(s->cfg.hbm_budget_bytes / 4);
    ATX_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    ATX_HIP(hipStreamCreateWithFlags(&s->adm_stream, hipStreamNonBlocking));
    ATX_HIP(hipEventCreate(&s->ev0));
    ATX_HIP(hipEventCreate(&s->ev1));
    ATX_HIP(hipEventCreate(&s->ev_prev));
    ATX_HIP(s->queue_ctr.reserve(256));
    ATX_HIP(s->counters.reserve(kCtrCount * sizeof(unsigned long long)));
    ATX_HIP(hipMemset(s->counters.get(), 0, kCtrCount * sizeof(unsigned long long)));
    s->grid = wavefront_full_grid(false);
    s->grid8 = wavefront_full_grid(false, 8);
    s->grid_flat = wavefront_full_grid(true);
    s->epoch = std::chrono::steady_clock::now();
    Server* raw = s.release();
    try {
        raw->retirer = std::thread(retirer_main, raw);
    } catch (...) {
        set_error("cannot start the server threads");
        delete raw;
        return ATX_ERR_STATE;
    }
    try {
        raw->worker = std::thread(worker_main, raw);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(raw->mu);
            raw->retire_stop = true;
        }
        raw->cv_retire.notify_all();
        raw->retirer.join();
        set_error("cannot start the server threads");
        delete raw;
        return ATX_ERR_STATE;
    }
    *out = reinterpret_cast<atx_server*>(raw);
    return ATX_OK;
}

int atx_server_destroy(atx_server* h)
{
    Server* s = cast(h);
    if (!s) return ATX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->retire_stop = true;   // the retire thread drains its queue first
    }
    s->cv_retire.notify_all();
    if (s->retirer.joinable()) s->retirer.join();
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->adm_stream) (void)hipStreamSynchronize(s->adm_stream);
    s->active.clear();
    s->queue.clear();
    s->all.clear();
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->ev_prev) (void)hipEventDestroy(s->ev_prev);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->adm_stream) (void)hipStreamDestroy(s->adm_stream);
    delete s;
    return ATX_OK;
}

int atx_server_submit(atx_server* h, const atx_scene_desc* d, uint64_t* request_id)
{
    Server* s = cast(h);
    ATX_CHECK_ARG(s && d, "NULL argument");
    ATX_CHECK_ARG(d->camera.width > 0 && d->camera.height > 0, "Width or height cannot be zero");
    ATX_CHECK_ARG((uint64_t)d->camera.width * d->camera.height < (1ull << 31), "viewport too large");
    ATX_CHECK_ARG(d->settings.max_bounces >= 0, "max_bounces must be >= 0");
    ATX_CHECK_ARG(d->spp >= 1, "spp must be >= 1");
    ATX_CHECK_ARG((uint64_t)d->camera.width * d->camera.height * std::min(d->spp, s->cfg.frames_per_tick) <
                      kMaxTickPaths,
                  "one tick of this request holds 2^31 or more paths; lower frames_per_tick");
    ATX_CHECK_ARG(d->num_spheres == 0 || d->spheres, "spheres is NULL");
    ATX_CHECK_ARG(d->num_materials == 0 || d->materials, "materials is NULL");
    ATX_CHECK_ARG(d->num_lights == 0 || d->lights, "lights is NULL");
    if (d->num_spheres > 0 && d->num_materials == 0) {
        set_error("scene has spheres but no material");
        return ATX_ERR_INVALID;
    }
    auto r = std::make_shared<Request>();
    r->spheres.assign(d->spheres, d->spheres + d->num_spheres);
    r->materials.assign(d->materials, d->materials + d->num_materials);
    r->lights.assign(d->lights, d->lights + d->num_lights);
    r->settings = d->settings;
    r->cam = d->camera;
    r->spp = d->spp;
    r->device_bytes = request_bytes(*r, s->cfg.kahan != 0, s->cfg.frames_per_tick);
    std::lock_guard<std::mutex> lk(s->mu);
    s->stats.submitted++;
    if (r->device_bytes > s->cfg.hbm_budget_bytes) {
        s->stats.rejected++;
        set_error("request needs %zu device bytes, budget is %llu", r->device_bytes,
                  (unsigned long long)s->cfg.hbm_budget_bytes);
        return ATX_ERR_BUDGET;
    }
    r->id = s->next_id++;
    r->t_submit = now_s(s);
    s->all[r->id] = r;
    s->queue.push_back(r);
    if (request_id) *request_id = r->id;
    s->cv_work.notify_one();
    return ATX_OK;
}

int atx_server_get_info(atx_server* h, uint64_t id, atx_request_info* info)
{
    Server* s = cast(h);
    ATX_CHECK_ARG(s && info, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) {
        set_error("unknown request %llu", (unsigned long long)id);
        return ATX_ERR_NOTFOUND;
    }
    const Request& r = *it->second;
    info->state = r.state;
    info->frames_done = r.frames_done;
    info->t_submit_s = r.t_submit;
    info->t_admit_s = r.t_admit;
    info->t_done_s = r.t_done;
    return ATX_OK;
}

int atx_server_wait(atx_server* h, uint64_t id, double timeout_s)
{
    Server* s = cast(h);
    ATX_CHECK_ARG(s, "server is NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) {
        set_error("unknown request %llu", (unsigned long long)id);
        return ATX_ERR_NOTFOUND;
    }
    std::shared_ptr<Request> r = it->second;
    auto done = [&r] { return r->state == ATX_REQ_DONE || r->state == ATX_REQ_FAILED; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        set_error("timed out waiting for request %llu", (unsigned long long)id);
        return ATX_ERR_STATE;
    }
    if (r->state == ATX_REQ_FAILED) {
        set_error("request %llu failed: %s", (unsigned long long)id, r->error.c_str());
        return ATX_ERR_STATE;
    }
    return ATX_OK;
}

int atx_server_get_result(atx_server* h, uint64_t id, uint32_t* rgba, uint64_t num_pixels, float* accum,
                          uint64_t num_floats, int release)
{
    Server* s = cast(h);
    ATX_CHECK_ARG(s, "server is NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) {
        set_error("unknown request %llu", (unsigned long long)id);
        return ATX_ERR_NOTFOUND;
    }
    Request& r = *it->second;
    if (r.state != ATX_REQ_DONE) {
        set_error("request %llu is not done", (unsigned long long)id);
        return ATX_ERR_STATE;
    }
    if (rgba) {
        ATX_CHECK_ARG(num_pixels >= r.rgba.size(), "rgba buffer too small");
        std::copy(r.rgba.begin(), r.rgba.end(), rgba);
    }
    if (accum) {
        ATX_CHECK_ARG(num_floats >= r.acc_host.size(), "accum buffer too small");
        std::copy(r.acc_host.begin(), r.acc_host.end(), accum);
    }
    if (release) s->all.erase(it);
    return ATX_OK;
}

int atx_server_get_stats(atx_server* h, atx_server_stats* st)
{
    Server* s = cast(h);
    ATX_CHECK_ARG(s && st, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    st->queued = s->queue.size();
    st->active = s->active.size();
    return ATX_OK;
}

}  // extern "C"
