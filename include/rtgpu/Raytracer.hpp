// Raytracer.hpp -- headless render entry with the shape of RaytracerBVHRenderer::Raytracer
// (RaytracerGPU_MastersProject/RaytracerBVH.hpp:510-571): construct with a scene function,
// then mainLoop() runs frames; each doIteration() re-uploads changed scene data, rebuilds the
// BVH on the device (S1), path traces raysPerPixel samples (S2), resolves and prints TIMINGS.
#pragma once

#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "Scene.hpp"

namespace RaytracerBVHRenderer {

using SceneFn = std::function<void(std::unique_ptr<RaytraceScene> &)>;

struct Timings {
    double updateSceneMs = 0, buildMs = 0, traceMs = 0, resolveMs = 0, frameMs = 0;
    unsigned long long segments = 0;
};

class Raytracer {
    rtgpu::Context context;
    std::unique_ptr<RaytraceScene> scene;
    uint32_t width, height;
    void *accum = nullptr, *resolved = nullptr, *segments = nullptr, *stream = nullptr;
    uint32_t iteration = 0;
    uint32_t randomState = 0;
    bool printTimings = true;
    Timings last;
    rt_params lastParams_{};
    std::vector<float> image_;

public:
    // The reference opens an 800x800 window and builds complexScene (RaytracerBVH.cpp:7-11, 511).
    explicit Raytracer(SceneFn sceneFn = complexScene, uint32_t width = 800, uint32_t height = 800, int device = 0);
    ~Raytracer();
    Raytracer(const Raytracer &) = delete;
    Raytracer &operator=(const Raytracer &) = delete;

    // the reference loops until the window closes; headless, it runs `frames` iterations
    void mainLoop(uint32_t frames = 1);
    void doIteration(float frameTime);

    // Config::RunRayPerPixelIncreasingDemo (Config.hpp:17-23, RaytracerBVH.hpp:522-570): frames from
    // `startRaysPerPixel`, raising the rays per pixel by `increaseAmount` every `runsBeforeIncrease` frames
    // until they exceed `maxRaysPerPixel`, then one line per level "index, mean frame time,\n" written to
    // `csvPath` (the reference's runtimes.csv; times in whole microseconds with the suffix "us", as a
    // std::chrono::microseconds prints).  Frame times are measured between successive loop starts, as
    // the reference's are.  Returns the number of levels written.
    struct SppSweep {
        uint32_t runsBeforeIncrease = 4, startRaysPerPixel = 100, maxRaysPerPixel = 200, increaseAmount = 5;
    };
    size_t sppSweep(const SppSweep &cfg, const std::string &csvPath = "runtimes.csv");

    void setRandomState(uint32_t s) { randomState = s; }
    void setPrintTimings(bool on) { printTimings = on; }
    RaytraceScene &getScene() { return *scene; }
    const Timings &timings() const { return last; }
    // RGBA32F accumulation of the last frame (W*H*4) and the resolved image
    std::vector<float> accumulation() const;
    const std::vector<float> &image() const { return image_; }
    void writePPM(const std::string &path) const;
    rt_params params() const;
    // the UBO of the last frame (including the random state it drew)
    const rt_params &lastParams() const { return lastParams_; }
    uint32_t imageWidth() const { return width; }
    uint32_t imageHeight() const { return height; }
};

}  // namespace RaytracerBVHRenderer
