// scene_dump.cpp -- rtgpu_scene_dump: run one of the Scenes functions against a host-only context
// (no GPU needed) and write the flattened scene plus its settings.
//
//   rtgpu_scene_dump SCENE OUTDIR [MODEL_DIR]      SCENE: complex|cornell|simple|mixed|spheres
#include <cstdio>
#include <map>
#include <string>

#include "dump.hpp"
#include "rtgpu/Scene.hpp"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s SCENE OUTDIR [MODEL_DIR]\n", argv[0]);
        return 2;
    }
    if (argc > 3) setModelDirectory(argv[3]);
    const std::map<std::string, void (*)(std::unique_ptr<RaytraceScene> &)> scenes = {
        {"complex", complexScene}, {"cornell", cornellBoxScene}, {"simple", simpleScene},
        {"mixed", cornellMixedScene}, {"spheres", randomSpheres}};
    auto it = scenes.find(argv[1]);
    if (it == scenes.end()) {
        std::fprintf(stderr, "unknown scene %s\n", argv[1]);
        return 2;
    }
    try {
        rtgpu::Context host(nullptr);
        auto scene = std::make_unique<RaytraceScene>(host);
        it->second(scene);
        rtgpu_dump::write_scene(argv[2], *scene);
        TransformComponent probe({1.0f, -2.0f, 3.0f}, {2.0f, 0.5f, 1.5f}, {0.3f, -1.1f, 0.7f});
        const rtgpu::mat4 pm = probe.mat4();
        rtgpu_dump::write_raw(std::string(argv[2]) + "/probe_mat4.bin", std::vector<float>(pm.data(), pm.data() + 16));
        std::printf("{\"scene\": \"%s\", \"triangles\": %u, \"spheres\": %u, \"materials\": %u, \"models\": %u, "
                    "\"depth\": %u, \"spp\": %u, \"fov\": %g}\n",
                    argv[1], scene->getTriangleCount(), scene->getSphereCount(), scene->getMaterialCount(),
                    scene->getModelCount(), scene->getMaxRaytraceDepth(), scene->getRaysPerPixel(),
                    (double)scene->getCamera().getVerticalFOV());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "rtgpu_scene_dump: %s\n", e.what());
        return 1;
    }
    return 0;
}
