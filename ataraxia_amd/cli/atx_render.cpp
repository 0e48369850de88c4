// atx_render -- headless driver: load scene.json, render frames, write a PPM.
//
//   atx_render <scene.json> <out.ppm> [width height frames [accumulation]]
//
// accumulation (1 or 0) overrides the file's setting.
// Uses the C++ façade exactly as the reference's editor drives its renderer
// once per displayed frame (Engine/src/main.cpp:211-220): onResize, then
// Camera::Resize, then Render; the image is written at the end instead of
// being uploaded to a Vulkan texture.
#include <cstdio>
#include <cstdlib>
#include <exception>

#include "ataraxia.hpp"

int main(int argc, char** argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <scene.json> <out.ppm> [width height frames [accumulation]]\n", argv[0]);
        return 2;
    }
    const uint32_t width = argc > 3 ? (uint32_t)std::strtoul(argv[3], nullptr, 10) : 640;
    const uint32_t height = argc > 4 ? (uint32_t)std::strtoul(argv[4], nullptr, 10) : 360;
    const uint32_t frames = argc > 5 ? (uint32_t)std::strtoul(argv[5], nullptr, 10) : 16;
    try {
        using namespace ataraxia;
        Scene scene = Utils::importScene(argv[1]);
        if (argc > 6) scene.settings.accumulation = std::strtoul(argv[6], nullptr, 10) != 0;
        Camera camera(scene.camera.getFov(), 0.1f, 100.0f, scene.camera.getPosition(), scene.camera.getDirection());
        Renderer renderer;
        renderer.setSettings(scene.settings);
        for (uint32_t f = 0; f < frames; ++f) {
            renderer.onResize(width, height);
            camera.Resize(width, height);
            renderer.Render(camera, scene);
        }
        const std::vector<uint32_t> img = renderer.getImage();
        check(atx_write_ppm(argv[2], img.data(), width, height, 1));
        std::printf("atx_render: %ux%u, %u frames, %llu scene upload(s) -> %s\n", width, height, frames,
                    (unsigned long long)renderer.sceneUploads(), argv[2]);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "atx_render: %s\n", e.what());
        return 1;
    }
    return 0;
}
