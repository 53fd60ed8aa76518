// main.cpp -- rtgpu_render: the program entry of the reference (RaytracerGPU_MastersProject/main.cpp:8-21,
// Config.hpp:6-23) without a window: render one of the Scenes functions for N frames and write a PPM.
//
//   rtgpu_render [--scene complex|cornell|simple|mixed|spheres] [--size WxH] [--frames N]
//                [--spp N] [--depth N] [--seed N] [--models DIR] [--out FILE.ppm] [--quiet]
//                [--spp-sweep RUNS,START,MAX,STEP [--csv FILE]]   (Config::RunRayPerPixelIncreasingDemo)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

#include <unistd.h>

#include "dump.hpp"
#include "rtgpu/Raytracer.hpp"

int main(int argc, char **argv) {
    std::string sceneName = "complex", out = "render.ppm";
    uint32_t w = 800, h = 800, frames = 1, seed = 0;
    int spp = -1, depth = -1;
    bool quiet = false;
    std::string dumpDir, csv = "runtimes.csv";
    bool sweep = false;
    RaytracerBVHRenderer::Raytracer::SppSweep sweepCfg;
    if (!std::getenv("RTGPU_MODEL_DIR")) {  // default: the models/ directory next to this executable
        char exe[4096];
        ssize_t n = readlink("/proc/self/exe", exe, sizeof exe - 1);
        if (n > 0) {
            exe[n] = 0;
            std::string dir(exe);
            setModelDirectory(dir.substr(0, dir.find_last_of('/')) + "/models");
        }
    }
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) { std::fprintf(stderr, "missing value for %s\n", a.c_str()); std::exit(2); }
            return argv[++i];
        };
        if (a == "--scene") sceneName = next();
        else if (a == "--size") { std::string s = next(); std::sscanf(s.c_str(), "%ux%u", &w, &h); }
        else if (a == "--frames") frames = (uint32_t)std::atoi(next().c_str());
        else if (a == "--spp") spp = std::atoi(next().c_str());
        else if (a == "--depth") depth = std::atoi(next().c_str());
        else if (a == "--seed") seed = (uint32_t)std::atoi(next().c_str());
        else if (a == "--models") setModelDirectory(next());
        else if (a == "--out") out = next();
        else if (a == "--quiet") quiet = true;
        else if (a == "--dump") dumpDir = next();
        else if (a == "--spp-sweep") {
            std::string s = next();
            sweep = std::sscanf(s.c_str(), "%u,%u,%u,%u", &sweepCfg.runsBeforeIncrease, &sweepCfg.startRaysPerPixel,
                                &sweepCfg.maxRaysPerPixel, &sweepCfg.increaseAmount) == 4;
            if (!sweep) { std::fprintf(stderr, "--spp-sweep wants RUNS,START,MAX,STEP\n"); return 2; }
        }
        else if (a == "--csv") csv = next();
        else { std::fprintf(stderr, "unknown argument %s\n", a.c_str()); return 2; }
    }
    const std::map<std::string, RaytracerBVHRenderer::SceneFn> scenes = {
        {"complex", complexScene}, {"cornell", cornellBoxScene}, {"simple", simpleScene},
        {"mixed", cornellMixedScene}, {"spheres", randomSpheres}};
    auto it = scenes.find(sceneName);
    if (it == scenes.end()) { std::fprintf(stderr, "unknown scene %s\n", sceneName.c_str()); return 2; }
    try {
        RaytracerBVHRenderer::Raytracer rt(it->second, w, h);
        if (spp >= 0) rt.getScene().setRaysPerPixel((uint32_t)spp);
        if (depth >= 0) rt.getScene().setMaxRaytraceDepth((uint32_t)depth);
        rt.setRandomState(seed);
        rt.setPrintTimings(!quiet);
        if (sweep) {
            const size_t levels = rt.sppSweep(sweepCfg, csv);
            std::printf("{\"spp_sweep_levels\": %zu, \"csv\": \"%s\"}\n", levels, csv.c_str());
        } else {
            rt.mainLoop(frames);
        }
        rt.writePPM(out);
        if (!dumpDir.empty()) {  // scene as flattened, the last frame's UBO and its accumulation image
            rtgpu_dump::write_scene(dumpDir, rt.getScene());
            rtgpu_dump::write_raw(dumpDir + "/params.bin", std::vector<rt_params>{rt.lastParams()});
            rtgpu_dump::write_raw(dumpDir + "/accum.bin", rt.accumulation());
        }
        const auto &t = rt.timings();
        std::printf("{\"scene\": \"%s\", \"width\": %u, \"height\": %u, \"spp\": %u, \"triangles\": %u, "
                    "\"spheres\": %u, \"trace_ms\": %.3f, \"segments\": %llu, \"out\": \"%s\"}\n",
                    sceneName.c_str(), w, h, rt.getScene().getRaysPerPixel(), rt.getScene().getTriangleCount(),
                    rt.getScene().getSphereCount(), t.traceMs, t.segments, out.c_str());
    } catch (const std::exception &e) {
        std::fprintf(stderr, "rtgpu_render: %s\n", e.what());
        return 1;
    }
    return 0;
}
