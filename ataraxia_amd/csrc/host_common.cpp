// host_common.cpp -- error state, scene layout, device queries, camera C ABI.
#include "host_common.h"

#include <stdlib.h>
#include <string.h>

#include "camera_host.h"

namespace atx {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

int build_scene_layout(const atx_sphere* spheres, uint64_t n, const atx_material* materials,
                       uint64_t m, const atx_light* lights, uint64_t l, HostSceneLayout* out)
{
    ATX_CHECK_ARG(n == 0 || spheres, "spheres is NULL with num_spheres=%llu", (unsigned long long)n);
    ATX_CHECK_ARG(m == 0 || materials, "materials is NULL with num_materials=%llu", (unsigned long long)m);
    ATX_CHECK_ARG(l == 0 || lights, "lights is NULL with num_lights=%llu", (unsigned long long)l);
    ATX_CHECK_ARG(n < (1ull << 31) && m < (1ull << 28) && l < (1ull << 28), "scene too large");
    if (n > 0 && m == 0) {
        set_error("scene has %llu spheres but no material", (unsigned long long)n);
        return ATX_ERR_INVALID;
    }
    out->sph.resize(4 * n);
    out->sph_mat.resize(n);
    for (uint64_t i = 0; i < n; ++i) {
        const atx_sphere& s = spheres[i];
        out->sph[4 * i + 0] = s.center[0];
        out->sph[4 * i + 1] = s.center[1];
        out->sph[4 * i + 2] = s.center[2];
        out->sph[4 * i + 3] = s.radius * s.radius;
        int32_t id = s.material;
        if ((uint32_t)id >= m) {
            fprintf(stderr, "Sphere ID out of bounds: %d\n", id);
            id = 0;
        }
        out->sph_mat[i] = id;
    }
    out->mat.resize(12 * m);
    for (uint64_t i = 0; i < m; ++i) {
        const atx_material& a = materials[i];
        float* o = &out->mat[12 * i];
        o[0] = a.albedo[0]; o[1] = a.albedo[1]; o[2] = a.albedo[2]; o[3] = a.roughness;
        o[4] = a.F0[0]; o[5] = a.F0[1]; o[6] = a.F0[2]; o[7] = a.metallic;
        o[8] = a.emission_color[0]; o[9] = a.emission_color[1]; o[10] = a.emission_color[2];
        o[11] = a.emission_intensity;
    }
    out->light.resize(8 * l);
    for (uint64_t i = 0; i < l; ++i) {
        const atx_light& a = lights[i];
        float* o = &out->light[8 * i];
        o[0] = a.position[0]; o[1] = a.position[1]; o[2] = a.position[2]; o[3] = a.intensity;
        o[4] = a.color[0]; o[5] = a.color[1]; o[6] = a.color[2]; o[7] = 0.0f;
    }
    return ATX_OK;
}

size_t scene_device_bytes(uint64_t n, uint64_t m, uint64_t l)
{
    return (16 * n + 16) + (4 * n + 16) + (48 * m + 16) + (32 * l + 16) + bvh_device_bytes(n);
}

int upload_scene(const atx_sphere* spheres, uint64_t n, const atx_material* materials, uint64_t m,
                 const atx_light* lights, uint64_t l, DevScene* out)
{
    HostSceneLayout L;
    int rc = build_scene_layout(spheres, n, materials, m, lights, l, &L);
    if (rc != ATX_OK) return rc;
    ATX_HIP(out->sph.reserve(L.sph.size() * 4 + 16));
    ATX_HIP(out->sph_mat.reserve(L.sph_mat.size() * 4 + 16));
    ATX_HIP(out->mat.reserve(L.mat.size() * 4 + 16));
    ATX_HIP(out->light.reserve(L.light.size() * 4 + 16));
    if (!L.sph.empty()) ATX_HIP(hipMemcpy(out->sph.get(), L.sph.data(), L.sph.size() * 4, hipMemcpyHostToDevice));
    if (!L.sph_mat.empty())
        ATX_HIP(hipMemcpy(out->sph_mat.get(), L.sph_mat.data(), L.sph_mat.size() * 4, hipMemcpyHostToDevice));
    if (!L.mat.empty()) ATX_HIP(hipMemcpy(out->mat.get(), L.mat.data(), L.mat.size() * 4, hipMemcpyHostToDevice));
    if (!L.light.empty())
        ATX_HIP(hipMemcpy(out->light.get(), L.light.data(), L.light.size() * 4, hipMemcpyHostToDevice));
    out->n_sph = (uint32_t)n;
    out->n_mat = (uint32_t)m;
    out->n_light = (uint32_t)l;

    HostBvh T;
    uint32_t layouts = kBvhLayouts;
    if (const char* env = getenv("ATX_BVH_LAYOUTS")) layouts = (uint32_t)atoi(env);
    build_bvh(L.sph.data(), (uint32_t)n, layouts, &T);
    out->bvh = BvhDev{};
    if (T.n_nodes > 0) {
        ATX_HIP(out->bvh_nodes.reserve(T.nodes.size() * 4 + 16));
        ATX_HIP(out->bvh_lsph.reserve(T.lsph.size() * 4 + 16));
        ATX_HIP(out->bvh_lidx.reserve(T.lidx.size() * 4 + 16));
        ATX_HIP(out->bvh_big.reserve(T.big.size() * 4 + 16));
        ATX_HIP(hipMemcpy(out->bvh_nodes.get(), T.nodes.data(), T.nodes.size() * 4, hipMemcpyHostToDevice));
        ATX_HIP(hipMemcpy(out->bvh_lsph.get(), T.lsph.data(), T.lsph.size() * 4, hipMemcpyHostToDevice));
        ATX_HIP(hipMemcpy(out->bvh_lidx.get(), T.lidx.data(), T.lidx.size() * 4, hipMemcpyHostToDevice));
        if (!T.big.empty())
            ATX_HIP(hipMemcpy(out->bvh_big.get(), T.big.data(), T.big.size() * 4, hipMemcpyHostToDevice));
        out->bvh = bvh_dev_from(T, out->bvh_nodes.get(), out->bvh_lsph.get(), out->bvh_lidx.get(),
                                out->bvh_big.get());
    }
    return ATX_OK;
}

BvhDev bvh_dev_from(const HostBvh& T, const void* nodes, const void* lsph, const void* lidx, const void* big)
{
    BvhDev b{};
    b.nodes = static_cast<const uint4*>(nodes);
    b.lsph = static_cast<const float4*>(lsph);
    b.lidx = static_cast<const int32_t*>(lidx);
    b.big = static_cast<const int32_t*>(big);
    b.n_nodes = T.n_nodes;
    b.n_big = (uint32_t)T.big.size();
    b.layout_mask = T.layouts - 1u;
    b.cx = T.cx;
    b.cy = T.cy;
    b.cz = T.cz;
    b.rb = T.rb;
    b.r_min = T.r_min;
    b.r_max = T.r_max;
    b.bx = T.base[0];
    b.by = T.base[1];
    b.bz = T.base[2];
    b.sx = T.scale[0];
    b.sy = T.scale[1];
    b.sz = T.scale[2];
    b.isx = T.inv_scale[0];
    b.isy = T.inv_scale[1];
    b.isz = T.inv_scale[2];
    b.ext = T.ext;
    return b;
}

}  // namespace atx

using namespace atx;

extern "C" {

const char* atx_last_error(void) { return g_last_error.c_str(); }

const char* atx_version(void) { return "ataraxia-amd 0.1 (gfx950)"; }

int atx_device_count(int* count)
{
    ATX_CHECK_ARG(count, "count is NULL");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *count = 0;
        set_error("hipGetDeviceCount: %s", hipGetErrorString(e));
        return ATX_ERR_NODEVICE;
    }
    *count = c;
    return ATX_OK;
}

int atx_device_synchronize(void)
{
    ATX_HIP(hipDeviceSynchronize());
    return ATX_OK;
}

int atx_camera_matrices(const atx_camera* cam, float* inv_proj, float* inv_view)
{
    ATX_CHECK_ARG(cam && inv_proj && inv_view, "NULL argument");
    ATX_CHECK_ARG(cam->width > 0 && cam->height > 0, "Width or height cannot be zero");
    camera_inverse_matrices(cam->position, cam->direction, cam->fov_deg, cam->near_clip, cam->far_clip,
                            cam->width, cam->height, inv_proj, inv_view);
    return ATX_OK;
}

int atx_camera_ray_directions(const atx_camera* cam, float* dirs, uint64_t num_floats)
{
    ATX_CHECK_ARG(cam && dirs, "NULL argument");
    ATX_CHECK_ARG(cam->width > 0 && cam->height > 0, "Width or height cannot be zero");
    ATX_CHECK_ARG(num_floats >= 3ull * cam->width * cam->height, "dirs holds %llu floats, need %llu",
                  (unsigned long long)num_floats, 3ull * cam->width * cam->height);
    float ip[16], iv[16];
    camera_inverse_matrices(cam->position, cam->direction, cam->fov_deg, cam->near_clip, cam->far_clip,
                            cam->width, cam->height, ip, iv);
    for (uint32_t y = 0; y < cam->height; ++y)
        for (uint32_t x = 0; x < cam->width; ++x) {
            const f3 d = camera_ray(ip, iv, x, y, cam->width, cam->height);
            float* o = dirs + 3 * ((size_t)x + (size_t)y * cam->width);
            o[0] = d.x; o[1] = d.y; o[2] = d.z;
        }
    return ATX_OK;
}

int atx_write_ppm(const char* path, const uint32_t* rgba, uint32_t width, uint32_t height, int flip_vertical)
{
    ATX_CHECK_ARG(path && rgba, "NULL argument");
    ATX_CHECK_ARG(width > 0 && height > 0, "Width or height cannot be zero");
    FILE* f = fopen(path, "wb");
    if (!f) {
        set_error("cannot open %s for writing", path);
        return ATX_ERR_IO;
    }
    fprintf(f, "P6\n%u %u\n255\n", width, height);
    std::vector<unsigned char> row((size_t)width * 3);
    bool ok = true;
    for (uint32_t r = 0; r < height && ok; ++r) {
        const uint32_t y = flip_vertical ? height - 1 - r : r;
        const uint32_t* src = rgba + (size_t)y * width;
        for (uint32_t x = 0; x < width; ++x) {
            row[3 * x + 0] = (unsigned char)(src[x] & 0xFF);
            row[3 * x + 1] = (unsigned char)((src[x] >> 8) & 0xFF);
            row[3 * x + 2] = (unsigned char)((src[x] >> 16) & 0xFF);
        }
        ok = fwrite(row.data(), 1, row.size(), f) == row.size();
    }
    if (fclose(f) != 0 || !ok) {
        set_error("write to %s failed", path);
        return ATX_ERR_IO;
    }
    return ATX_OK;
}

}  // extern "C"
