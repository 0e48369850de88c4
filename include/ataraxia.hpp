// ataraxia.hpp -- header-only C++17 façade over the C ABI (atx.h) that keeps
// the reference's class and method names, so that code written against the
// reference's Renderer / Scene / SceneNode / Camera / Utils compiles against
// libatx with the GUI parts removed.
//
// Reference interfaces mirrored (paths relative to the reference repository):
//   Renderer   Engine/include/Renderer.h:18-31   (getImage returns RGBA8 pixels instead of a Vulkan Image)
//   Camera     Engine/include/Camera.h:14-58     (no interactive onUpdate)
//   SceneNode  Engine/include/SceneNode.h:23-66
//   Scene      Engine/include/Scene.h:58-80
//   Utils      Engine/include/Utils.h:11-23      (importScene / exportScene)
// Errors from the library are thrown as ataraxia::Error (the reference exits
// the process on CUDA errors, DeviceMemory.h:7-16).
#pragma once

#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "atx.h"

namespace ataraxia {

struct Error : std::runtime_error {
    int code;
    Error(int c, const char* msg) : std::runtime_error(msg ? msg : "atx error"), code(c) {}
};

inline void check(int rc)
{
    if (rc != ATX_OK) throw Error(rc, atx_last_error());
}

using vec3 = std::array<float, 3>;
using quat = std::array<float, 4>;   // x, y, z, w
using Sphere = atx_sphere;
using Material = atx_material;
using Light = atx_light;

struct Settings {   // Scene.h:49-56
    bool accumulation = true;
    bool skyLight = false;
    int maxBounces = 15;
    atx_settings c() const
    {
        atx_settings s{};
        s.accumulation = accumulation ? 1 : 0;
        s.sky_light = skyLight ? 1 : 0;
        s.max_bounces = maxBounces;
        return s;
    }
    static Settings from(const atx_settings& s) { return Settings{s.accumulation != 0, s.sky_light != 0, s.max_bounces}; }
};

inline Material makeMaterial(const vec3& albedo, float roughness, float metallic, const vec3& emissionColor,
                             float emissionIntensity, int id)   // Material(...) of Scene.h:37-40
{
    Material m{};
    for (int i = 0; i < 3; ++i) {
        m.albedo[i] = albedo[i];
        m.F0[i] = 0.04f;
        m.emission_color[i] = emissionColor[i];
    }
    m.roughness = roughness;
    m.metallic = metallic;
    m.emission_intensity = emissionIntensity;
    m.id = id;
    return m;
}

inline Light makeLight(const vec3& pos, const vec3& col, float intensity)
{
    Light l{};
    for (int i = 0; i < 3; ++i) {
        l.position[i] = pos[i];
        l.color[i] = col[i];
    }
    l.intensity = intensity;
    return l;
}

inline Sphere makeSphere(const vec3& c, float r, int materialId)
{
    Sphere s{};
    s.center[0] = c[0];
    s.center[1] = c[1];
    s.center[2] = c[2];
    s.radius = r;
    s.material = materialId;
    return s;
}

class Camera {
public:
    Camera() = default;
    Camera(float fov, float nearClip, float farClip)   // Camera.cpp:14-21
        : m_fov(fov), m_near(nearClip), m_far(farClip), m_position{0.0f, 0.0f, 3.0f}, m_direction{0.0f, 0.0f, -1.0f}
    {
    }
    Camera(float fov, float nearClip, float farClip, vec3 position, vec3 direction)
        : m_fov(fov), m_near(nearClip), m_far(farClip), m_position(position), m_direction(direction)
    {
    }
    void Resize(uint32_t width, uint32_t height)   // Camera.cpp:110-127
    {
        if (width == 0 || height == 0) {
            std::fprintf(stderr, "Error: Width or height cannot be zero.\n");
            return;
        }
        m_width = width;
        m_height = height;
    }
    const vec3& getPosition() const { return m_position; }
    const vec3& getDirection() const { return m_direction; }
    const float& getFov() const { return m_fov; }
    void setPosition(const vec3& p) { m_position = p; }
    void setDirection(const vec3& d) { m_direction = d; }
    void setFov(float fov) { m_fov = fov; }
    uint32_t getWidth() const { return m_width; }
    uint32_t getHeight() const { return m_height; }
    // Camera::UpdateRayDirection's result, computed by the library (3 floats per pixel)
    std::vector<float> getRayDirection() const
    {
        std::vector<float> d((size_t)m_width * m_height * 3);
        const atx_camera c = desc();
        check(atx_camera_ray_directions(&c, d.data(), d.size()));
        return d;
    }
    atx_camera desc() const
    {
        atx_camera c{};
        for (int i = 0; i < 3; ++i) {
            c.position[i] = m_position[i];
            c.direction[i] = m_direction[i];
        }
        c.fov_deg = m_fov;
        c.near_clip = m_near;
        c.far_clip = m_far;
        c.width = m_width;
        c.height = m_height;
        return c;
    }

private:
    float m_fov = 45.0f, m_near = 0.1f, m_far = 100.0f;
    vec3 m_position{0.0f, 0.0f, 0.0f}, m_direction{0.0f, 0.0f, 0.0f};
    uint32_t m_width = 1600, m_height = 900;
};

class SceneNode {
public:
    SceneNode() : m_name("Untitled") {}
    explicit SceneNode(const std::string& name) : m_name(name) {}
    void setPosition(const vec3& p) { m_position = p; }
    void setRotation(const quat& q) { m_rotation = q; }
    void setScale(const vec3& s) { m_scale = s; }
    const vec3& getPosition() const { return m_position; }
    const quat& getRotation() const { return m_rotation; }
    const vec3& getScale() const { return m_scale; }
    void addChild(std::shared_ptr<SceneNode> child) { m_children.push_back(std::move(child)); }
    void removeChild(const std::shared_ptr<SceneNode>& child)
    {
        std::vector<std::shared_ptr<SceneNode>> keep;
        for (auto& c : m_children)
            if (c != child) keep.push_back(c);
        m_children.swap(keep);
    }
    const std::vector<std::shared_ptr<SceneNode>>& getChildren() const { return m_children; }
    void addSphere(const Sphere& s) { m_spheres.push_back(s); }
    void removeSphere(int i)
    {
        if (i >= 0 && i < (int)m_spheres.size()) m_spheres.erase(m_spheres.begin() + i);
    }
    const std::vector<Sphere>& getSpheres() const { return m_spheres; }
    const std::string& getName() const { return m_name; }
    void setName(const std::string& n) { m_name = n; }

private:
    std::string m_name;
    vec3 m_position{0.0f, 0.0f, 0.0f};
    quat m_rotation{0.0f, 0.0f, 0.0f, 0.0f};   // value-initialised glm::quat
    vec3 m_scale{1.0f, 1.0f, 1.0f};
    std::vector<std::shared_ptr<SceneNode>> m_children;
    std::vector<Sphere> m_spheres;
};

struct Scene {   // Scene.h:58-80
    std::shared_ptr<SceneNode> rootNode = std::make_shared<SceneNode>("Scene");
    std::vector<Material> materials;
    std::vector<Light> lights;
    Settings settings;
    Camera camera;
};

namespace detail {

struct GraphHandle {
    atx_scene_graph* g = nullptr;
    GraphHandle() { check(atx_scene_create(&g)); }
    explicit GraphHandle(atx_scene_graph* h) : g(h) {}
    ~GraphHandle() { atx_scene_destroy(g); }
    GraphHandle(const GraphHandle&) = delete;
    GraphHandle& operator=(const GraphHandle&) = delete;
};

inline void add_subtree(atx_scene_graph* g, uint32_t id, const SceneNode& n)
{
    for (const Sphere& s : n.getSpheres()) check(atx_scene_add_sphere(g, id, &s));
    for (const auto& c : n.getChildren()) {
        uint32_t k = 0;
        check(atx_scene_add_node(g, id, c->getName().c_str(), c->getPosition().data(), c->getRotation().data(),
                                 c->getScale().data(), &k));
        add_subtree(g, k, *c);
    }
}

// The C++ scene as a library scene graph.
inline void to_graph(const Scene& sc, atx_scene_graph* g)
{
    const SceneNode& r = *sc.rootNode;
    check(atx_scene_set_node_transform(g, 0, r.getPosition().data(), r.getRotation().data(), r.getScale().data()));
    add_subtree(g, 0, r);
    for (const Material& m : sc.materials) check(atx_scene_add_material(g, &m));
    for (const Light& l : sc.lights) check(atx_scene_add_light(g, &l));
    const atx_settings s = sc.settings.c();
    check(atx_scene_set_settings(g, &s));
    check(atx_scene_set_camera(g, sc.camera.getPosition().data(), sc.camera.getDirection().data(),
                               sc.camera.getFov()));
}

// Renderer::traverseSceneGraph through the library (Renderer.cu:67-96).
inline std::vector<Sphere> flatten(const Scene& sc)
{
    GraphHandle h;
    to_graph(sc, h.g);
    uint64_t n = 0;
    check(atx_scene_flatten(h.g, nullptr, 0, &n));
    std::vector<Sphere> out(n);
    if (n) check(atx_scene_flatten(h.g, out.data(), n, &n));
    return out;
}

}  // namespace detail

namespace Utils {

// Utils::importScene (Utils.cpp:175-187): a missing file gives Scene().
inline Scene importScene(const std::string& filename)
{
    atx_scene_graph* raw = nullptr;
    check(atx_scene_load_json(filename.c_str(), &raw));
    detail::GraphHandle h(raw);
    Scene sc;
    uint64_t nodes = 0, nm = 0, nl = 0;
    check(atx_scene_counts(h.g, &nodes, nullptr, nullptr, nullptr));
    std::vector<std::shared_ptr<SceneNode>> byId(nodes);
    for (uint64_t i = 0; i < nodes; ++i) {
        char name[256];
        vec3 p{}, sc3{};
        quat q{};
        uint32_t parent = 0;
        uint64_t ns = 0;
        check(atx_scene_get_node(h.g, (uint32_t)i, name, sizeof(name), p.data(), q.data(), sc3.data(), &parent, &ns));
        auto n = i == 0 ? sc.rootNode : std::make_shared<SceneNode>(name);
        n->setName(name);
        n->setPosition(p);
        n->setRotation(q);
        n->setScale(sc3);
        std::vector<Sphere> sph(ns);
        if (ns) check(atx_scene_get_node_spheres(h.g, (uint32_t)i, sph.data(), ns, &ns));
        for (const Sphere& x : sph) n->addSphere(x);
        byId[i] = n;
        if (i > 0) byId[parent]->addChild(n);   // nodes are numbered parents first
    }
    check(atx_scene_get_materials(h.g, nullptr, 0, &nm));
    sc.materials.resize(nm);
    if (nm) check(atx_scene_get_materials(h.g, sc.materials.data(), nm, &nm));
    check(atx_scene_get_lights(h.g, nullptr, 0, &nl));
    sc.lights.resize(nl);
    if (nl) check(atx_scene_get_lights(h.g, sc.lights.data(), nl, &nl));
    atx_settings s{};
    check(atx_scene_get_settings(h.g, &s));
    sc.settings = Settings::from(s);
    vec3 p{}, d{};
    float fov = 45.0f;
    check(atx_scene_get_camera(h.g, p.data(), d.data(), &fov));
    sc.camera.setPosition(p);
    sc.camera.setDirection(d);
    sc.camera.setFov(fov);
    return sc;
}

// Utils::exportScene (Utils.cpp:53-62)
inline void exportScene(const Scene& scene, const std::string& filename)
{
    detail::GraphHandle h;
    detail::to_graph(scene, h.g);
    check(atx_scene_save_json(h.g, filename.c_str()));
}

}  // namespace Utils

class Renderer {
public:
    Renderer() { check(atx_renderer_create(&m_r)); }
    ~Renderer() { atx_renderer_destroy(m_r); }
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    void onResize(uint32_t width, uint32_t height)   // Renderer.cu:98-146
    {
        check(atx_renderer_on_resize(m_r, width, height));
        m_width = width;
        m_height = height;
    }
    // Renderer.cu:173-249: the reference re-uploads the scene when the Scene
    // object changes or the frame index is 1 (every frame with accumulation
    // off).  Here the scene is flattened at those points as in the reference,
    // and uploaded only when the flattened spheres, materials or lights differ
    // from what the device holds: the device then holds exactly what the
    // reference's upload would give, without rebuilding the sphere tree and
    // the first-bounce cache every frame.
    void Render(Camera& camera, const Scene& scene)
    {
        uint32_t fi = 0;
        check(atx_renderer_get_frame_index(m_r, &fi));
        if (m_scene != &scene || fi == 1) {
            m_scene = &scene;
            std::vector<Sphere> sph = detail::flatten(scene);
            if (!m_uploaded || !same(sph, m_sph) || !same(scene.materials, m_mat) || !same(scene.lights, m_lig)) {
                check(atx_renderer_set_scene(m_r, sph.data(), sph.size(), scene.materials.data(),
                                             scene.materials.size(), scene.lights.data(), scene.lights.size()));
                m_sph = std::move(sph);
                m_mat = scene.materials;
                m_lig = scene.lights;
                m_uploaded = true;
                ++m_uploads;
            }
        }
        camera.Resize(m_width, m_height);
        const atx_camera c = camera.desc();
        check(atx_renderer_render(m_r, &c));
    }
    // Scene uploads so far (each one rebuilds the device scene and its tree).
    uint64_t sceneUploads() const { return m_uploads; }
    // The RGBA8 pixels the reference uploads to its Vulkan Image.
    std::vector<uint32_t> getImage() const
    {
        std::vector<uint32_t> px((size_t)m_width * m_height);
        check(atx_renderer_get_image(m_r, px.data(), px.size()));
        return px;
    }
    Settings getSettings() const
    {
        atx_settings s{};
        check(atx_renderer_get_settings(m_r, &s));
        return Settings::from(s);
    }
    void setSettings(const Settings& s)
    {
        const atx_settings c = s.c();
        check(atx_renderer_set_settings(m_r, &c));
    }
    void resetFrameIndex() { check(atx_renderer_reset_frame_index(m_r)); }
    atx_renderer* handle() { return m_r; }

private:
    template <class T>
    static bool same(const std::vector<T>& a, const std::vector<T>& b)
    {
        return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(T)) == 0);
    }
    atx_renderer* m_r = nullptr;
    const Scene* m_scene = nullptr;
    std::vector<Sphere> m_sph;
    std::vector<Material> m_mat;
    std::vector<Light> m_lig;
    bool m_uploaded = false;
    uint64_t m_uploads = 0;
    uint32_t m_width = 0, m_height = 0;
};

}  // namespace ataraxia
