// host_common.h -- error reporting and device-buffer ownership for libatx.
//
// The reference exits the process on any CUDA error (DeviceMemory.h:7-16);
// libatx instead records a thread-local message and returns a status code.
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/atx.h"
#include "bvh.h"

namespace atx {

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));

#define ATX_HIP(call)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            ::atx::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),         \
                             __FILE__, __LINE__);                                          \
            return e_ == hipErrorOutOfMemory ? ATX_ERR_NOMEM : ATX_ERR_HIP;                \
        }                                                                                  \
    } while (0)

#define ATX_CHECK_ARG(cond, ...)          \
    do {                                  \
        if (!(cond)) {                    \
            ::atx::set_error(__VA_ARGS__); \
            return ATX_ERR_INVALID;       \
        }                                 \
    } while (0)

// Device blocks kept for reuse.  hipFree waits for the device and costs
// ~70 us per call even when it is idle, so an owner that allocates and
// frees per request (the server) returns blocks here and takes them back
// for later requests.  A block is handed out for a request of at least half
// its size; past `limit` pooled bytes the largest blocks are freed.  Blocks
// are only ever given back once the device no longer uses them.
class DevPool {
public:
    explicit DevPool(size_t limit = 0) : limit_(limit) {}
    ~DevPool() { trim(0); }
    DevPool(const DevPool&) = delete;
    DevPool& operator=(const DevPool&) = delete;
    void set_limit(size_t limit) { std::lock_guard<std::mutex> lk(mu_); limit_ = limit; }
    // Capacity beyond the bytes asked for that take() may still hand out, in
    // total over the next calls (SIZE_MAX: no limit).  An owner that meters
    // device bytes against a budget sets it to what the budget has left.
    void set_slack(size_t slack) { std::lock_guard<std::mutex> lk(mu_); slack_ = slack; }
    // A pooled block of capacity in [bytes, 2 * bytes + 64 KiB] whose excess
    // over `bytes` fits the slack, or nullptr.
    void* take(size_t bytes, size_t* cap)
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = free_.lower_bound(bytes);
        if (it == free_.end() || it->first > 2 * bytes + (64u << 10) || it->first - bytes > slack_) return nullptr;
        if (slack_ != SIZE_MAX) slack_ -= it->first - bytes;
        void* p = it->second;
        *cap = it->first;
        pooled_ -= it->first;
        free_.erase(it);
        return p;
    }
    void give(void* p, size_t cap)
    {
        std::lock_guard<std::mutex> lk(mu_);
        free_.emplace(cap, p);
        pooled_ += cap;
        while (pooled_ > limit_ && !free_.empty()) {
            auto big = std::prev(free_.end());
            pooled_ -= big->first;
            (void)hipFree(big->second);
            free_.erase(big);
        }
    }
    void trim(size_t keep)
    {
        std::lock_guard<std::mutex> lk(mu_);
        while (pooled_ > keep && !free_.empty()) {
            auto big = std::prev(free_.end());
            pooled_ -= big->first;
            (void)hipFree(big->second);
            free_.erase(big);
        }
    }
    size_t pooled() const { std::lock_guard<std::mutex> lk(mu_); return pooled_; }

private:
    mutable std::mutex mu_;
    std::multimap<size_t, void*> free_;   // capacity -> block
    size_t pooled_ = 0, limit_ = 0, slack_ = SIZE_MAX;
};

// Move-only owner of one device allocation (the reference's CudaBuffer has a
// defaulted copy constructor, DeviceMemory.h:33, which can double-free).
// With a pool (use_pool), blocks come from and go back to it.
class DevBuf {
public:
    DevBuf() = default;
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_), pool_(o.pool_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p_ = o.p_; n_ = o.n_; pool_ = o.pool_;
            o.p_ = nullptr; o.n_ = 0;
        }
        return *this;
    }
    void use_pool(DevPool* pool) { pool_ = pool; }
    // Grows to at least `bytes` (contents are not preserved on growth).
    hipError_t reserve(size_t bytes)
    {
        if (bytes <= n_ && p_) return hipSuccess;
        release();
        if (bytes == 0) return hipSuccess;
        if (pool_) {
            size_t cap = 0;
            if (void* q = pool_->take(bytes, &cap)) {
                p_ = q;
                n_ = cap;
                return hipSuccess;
            }
        }
        hipError_t e = hipMalloc(&p_, bytes);
        if (e != hipSuccess) { p_ = nullptr; return e; }
        n_ = bytes;
        return hipSuccess;
    }
    void release()
    {
        if (p_) {
            if (pool_) pool_->give(p_, n_);
            else (void)hipFree(p_);
        }
        p_ = nullptr;
        n_ = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p_); }
    void* get() const { return p_; }
    size_t bytes() const { return n_; }

private:
    void* p_ = nullptr;
    size_t n_ = 0;
    DevPool* pool_ = nullptr;
};

// Device-layout copy of one flattened scene (see device_types.h).
struct HostSceneLayout {
    std::vector<float> sph;       // 4 per sphere: cx, cy, cz, r*r
    std::vector<int32_t> sph_mat; // material index per sphere (clamped)
    std::vector<float> mat;       // 12 per material
    std::vector<float> light;     // 8 per light
};

// Builds the device layout; ids outside [0, num_materials) are clamped to 0
// with a message on stderr, as Renderer::allocateDeviceMemory does
// (Engine/src/Renderer.cu:30-37).  Fails if spheres exist but no material.
int build_scene_layout(const atx_sphere* spheres, uint64_t n, const atx_material* materials,
                       uint64_t m, const atx_light* lights, uint64_t l, HostSceneLayout* out);

// One scene's geometry in HBM (device layout of device_types.h), with its
// sphere BVH (bvh.h; bvh.n_nodes = 0 when the scene has no tree).
struct DevScene {
    DevBuf sph, sph_mat, mat, light;
    DevBuf bvh_nodes, bvh_lsph, bvh_lidx, bvh_big;
    uint32_t n_sph = 0, n_mat = 0, n_light = 0;
    BvhDev bvh{};
    void use_pool(DevPool* pool)
    {
        for (DevBuf* b : {&sph, &sph_mat, &mat, &light, &bvh_nodes, &bvh_lsph, &bvh_lidx, &bvh_big})
            b->use_pool(pool);
    }
    size_t bytes() const
    {
        return sph.bytes() + sph_mat.bytes() + mat.bytes() + light.bytes() + bvh_nodes.bytes() +
               bvh_lsph.bytes() + bvh_lidx.bytes() + bvh_big.bytes();
    }
    // The tree to hand to a kernel: the scene's, or none for the flat loop.
    BvhDev accel(int mode) const
    {
        BvhDev b{};
        if (mode == ATX_ACCEL_BVH) b = bvh;
        return b;
    }
};

// Validates, converts and uploads (synchronously) a flattened scene.
int upload_scene(const atx_sphere* spheres, uint64_t n, const atx_material* materials, uint64_t m,
                 const atx_light* lights, uint64_t l, DevScene* out);

// Kernel view of a host-built tree whose arrays were copied to the given
// device addresses (HostBvh::nodes, lsph, lidx and big, in that order).
BvhDev bvh_dev_from(const HostBvh& T, const void* nodes, const void* lsph, const void* lidx, const void* big);

// Device bytes a scene of this size occupies (geometry and tree).
size_t scene_device_bytes(uint64_t n, uint64_t m, uint64_t l);

}  // namespace atx
