// render_noalloc.hip -- the allocation-free render entry point.
//
// atx_render_frames_noalloc renders frames of one scene with the plain kernel
// (render_plain.hip) exactly as atx_renderer_render_frames does with
// ATX_KERNEL_PLAIN, but allocates no device memory: the caller passes the
// accumulation, compensation and image buffers, and everything else the
// kernel reads (the scene in its device layout, the sphere tree, the
// statistics counters) lives in one file-scope device array, `scratch`, of
// kScratchBytes.  It is for callers that cannot call hipMalloc (serving
// loops, harnesses that forbid allocation).
//
// Each device holds its own copy of `scratch` for the life of the process,
// at its own address; that address is looked up once per device with
// hipGetSymbolAddress and only the looked-up pointer is handed to kernels and
// copies (the symbol's host-side name is not a device address).  Every call
// on a device shares its array: calls that overlap on one device (two host
// threads, or two streams) race on it.  A call returns after its frames are
// rendered, so consecutive calls on one thread do not overlap.
#include <mutex>
#include <string.h>

#include "camera_host.h"
#include "device_types.h"
#include "host_common.h"
#include "wavefront.h"

namespace atx {
namespace {

constexpr size_t kScratchBytes = 8u << 20;   // 8 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's `scratch`, looked up on the first call
// made with the device current.
int scratch_address(int dev, unsigned char** out)
{
    ATX_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        ATX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ATX_CHECK_ARG(p != nullptr, "scratch has no device address");
        g_scratch_ptr[dev] = p;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[dev]);
    return ATX_OK;
}

size_t align256(size_t v) { return (v + 255u) & ~(size_t)255u; }

}  // namespace
}  // namespace atx

using namespace atx;

extern "C" int atx_render_frames_noalloc(const atx_sphere* spheres, uint64_t num_spheres,
                                         const atx_material* materials, uint64_t num_materials,
                                         const atx_light* lights, uint64_t num_lights, const atx_settings* settings,
                                         const atx_camera* cam, uint32_t first_frame, uint32_t frames, int accel,
                                         int accumulator, float* accum, float* comp, uint32_t* image,
                                         void* stream_handle)
{
    ATX_CHECK_ARG(settings && cam, "settings or camera is NULL");
    ATX_CHECK_ARG(cam->width > 0 && cam->height > 0, "Width or height cannot be zero");
    ATX_CHECK_ARG(settings->max_bounces >= 0, "max_bounces must be >= 0");
    ATX_CHECK_ARG(first_frame >= 1, "first_frame must be >= 1 (the reference's frame index starts at 1)");
    ATX_CHECK_ARG(accel == ATX_ACCEL_BVH || accel == ATX_ACCEL_NONE, "unknown accel %d", accel);
    ATX_CHECK_ARG(accumulator == ATX_ACCUM_NAIVE || accumulator == ATX_ACCUM_KAHAN, "unknown accumulator %d",
                  accumulator);
    ATX_CHECK_ARG(accum && image, "accum and image must be device buffers of width*height entries");
    ATX_CHECK_ARG(accumulator == ATX_ACCUM_NAIVE || comp, "Kahan accumulation needs the comp buffer");
    if (frames == 0) return ATX_OK;
    ATX_CHECK_ARG(!settings->accumulation || first_frame + (frames - 1u) >= first_frame, "frame index overflows");

    HostSceneLayout L;
    int rc = build_scene_layout(spheres, num_spheres, materials, num_materials, lights, num_lights, &L);
    if (rc != ATX_OK) return rc;
    HostBvh T;
    if (accel == ATX_ACCEL_BVH) build_bvh(L.sph.data(), (uint32_t)num_spheres, kBvhLayouts, &T);

    // scratch layout: counters | spheres | material ids | materials | lights | tree
    const size_t off_ctr = 0;
    const size_t off_sph = align256(off_ctr + kCtrCount * sizeof(unsigned long long));
    const size_t off_smat = align256(off_sph + L.sph.size() * 4);
    const size_t off_mat = align256(off_smat + L.sph_mat.size() * 4);
    const size_t off_light = align256(off_mat + L.mat.size() * 4);
    const size_t off_nodes = align256(off_light + L.light.size() * 4);
    const size_t off_lsph = align256(off_nodes + T.nodes.size() * 4);
    const size_t off_lidx = align256(off_lsph + T.lsph.size() * 4);
    const size_t off_big = align256(off_lidx + T.lidx.size() * 4);
    const size_t total = align256(off_big + T.big.size() * 4);
    if (total > kScratchBytes) {
        set_error("scene needs %zu bytes of device scratch; atx_render_frames_noalloc holds at most %zu "
                  "(use atx_renderer_* for larger scenes)",
                  total, kScratchBytes);
        return ATX_ERR_INVALID;
    }

    int dev = 0;
    ATX_HIP(hipGetDevice(&dev));
    unsigned char* base = nullptr;
    rc = scratch_address(dev, &base);
    if (rc != ATX_OK) return rc;
    hipStream_t stream = static_cast<hipStream_t>(stream_handle);

    auto put = [&](size_t off, const void* src, size_t bytes) -> hipError_t {
        return bytes ? hipMemcpyAsync(base + off, src, bytes, hipMemcpyHostToDevice, stream) : hipSuccess;
    };
    ATX_HIP(put(off_sph, L.sph.data(), L.sph.size() * 4));
    ATX_HIP(put(off_smat, L.sph_mat.data(), L.sph_mat.size() * 4));
    ATX_HIP(put(off_mat, L.mat.data(), L.mat.size() * 4));
    ATX_HIP(put(off_light, L.light.data(), L.light.size() * 4));
    ATX_HIP(put(off_nodes, T.nodes.data(), T.nodes.size() * 4));
    ATX_HIP(put(off_lsph, T.lsph.data(), T.lsph.size() * 4));
    ATX_HIP(put(off_lidx, T.lidx.data(), T.lidx.size() * 4));
    ATX_HIP(put(off_big, T.big.data(), T.big.size() * 4));
    // the counters start at zero in every call (a fill kernel, not a memset node)
    ATX_HIP(launch_zero_u32(reinterpret_cast<unsigned int*>(base + off_ctr),
                            (uint32_t)(kCtrCount * sizeof(unsigned long long) / 4), stream));

    PlainArgs a{};
    a.sc.sph = reinterpret_cast<const float4*>(base + off_sph);
    a.sc.sph_mat = reinterpret_cast<const int32_t*>(base + off_smat);
    a.sc.mat = reinterpret_cast<const float4*>(base + off_mat);
    a.sc.light = reinterpret_cast<const float4*>(base + off_light);
    a.sc.n_sph = (uint32_t)num_spheres;
    a.sc.n_mat = (uint32_t)num_materials;
    a.sc.n_light = (uint32_t)num_lights;
    if (T.n_nodes > 0)
        a.sc.bvh = bvh_dev_from(T, base + off_nodes, base + off_lsph, base + off_lidx, base + off_big);
    camera_inverse_matrices(cam->position, cam->direction, cam->fov_deg, cam->near_clip, cam->far_clip,
                            cam->width, cam->height, a.cam.inv_proj, a.cam.inv_view);
    memcpy(a.cam.pos, cam->position, sizeof(a.cam.pos));
    a.cam.width = cam->width;
    a.cam.height = cam->height;
    a.max_bounces = settings->max_bounces;
    a.sky_light = settings->sky_light ? 1u : 0u;
    a.accum = reinterpret_cast<float4*>(accum);
    a.comp = accumulator == ATX_ACCUM_KAHAN ? reinterpret_cast<float4*>(comp) : nullptr;
    a.image = image;
    a.counters = reinterpret_cast<unsigned long long*>(base + off_ctr);

    // the variant atx_renderer_render_frames picks for the plain kernel
    uint32_t variant = 0;
    if ((size_t)num_spheres * sizeof(float4) <= 64u * 1024u) variant |= kVarLds;
    if (accumulator == ATX_ACCUM_KAHAN) variant |= kVarKahan;
    if (a.sc.bvh.n_nodes) variant = (variant & ~(uint32_t)kVarLds) | kVarBvh;

    uint32_t fi = first_frame;
    for (uint32_t f = 0; f < frames; ++f) {
        a.frame = fi;
        ATX_HIP(launch_plain(a, variant, stream));
        fi = settings->accumulation ? fi + 1u : 1u;   // Renderer.cu:245-248
    }
    // the host copies above read L and T, which end with this call
    ATX_HIP(hipStreamSynchronize(stream));
    return ATX_OK;
}
