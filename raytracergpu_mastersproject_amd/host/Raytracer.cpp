// Raytracer.cpp -- headless render entry with the shape of RaytracerBVHRenderer::Raytracer
// (RaytracerGPU_MastersProject/RaytracerBVH.hpp:206-571, RaytracerBVH.cpp:734-1078).
#include "rtgpu/Raytracer.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <fstream>
#include <random>

namespace RaytracerBVHRenderer {

namespace {
using clock_t_ = std::chrono::high_resolution_clock;
double ms_since(clock_t_::time_point t0) {
    return std::chrono::duration<double, std::milli>(clock_t_::now() - t0).count();
}
}  // namespace

Raytracer::Raytracer(SceneFn sceneFn, uint32_t w, uint32_t h, int device) : context(device), width(w), height(h) {
    if (w == 0 || h == 0) throw std::invalid_argument("image size must be positive");
    scene = std::make_unique<RaytraceScene>(context);
    sceneFn(scene);  // the scene functions end with prepForRender()
    const size_t npix = (size_t)w * h;
    context.check(rt_util_malloc(&accum, 16 * npix), "accumulation image");
    context.check(rt_util_malloc(&resolved, 16 * npix), "resolved image");
    context.check(rt_util_malloc(&segments, 8), "segment counter");
    context.check(rt_util_stream_create(&stream), "stream");
}

Raytracer::~Raytracer() {
    if (stream) rt_util_stream_synchronize(stream);
    for (void *p : {accum, resolved, segments})
        if (p) rt_util_free(p);
    if (stream) rt_util_stream_destroy(stream);
}

rt_params Raytracer::params() const {
    // RaytracerBVH.hpp:340-352 fills the UBO: camera pose, FOV, counts, depth and a random state
    rt_params p{};
    const CameraGameObject &cam = const_cast<RaytraceScene &>(*scene).getCamera();
    const rtgpu::vec3 pos = cam.position, at = cam.lookAt, up = cam.up;
    p.camPos[0] = pos.x; p.camPos[1] = pos.y; p.camPos[2] = pos.z;
    p.camLookAt[0] = at.x; p.camLookAt[1] = at.y; p.camLookAt[2] = at.z;
    p.camUpDir[0] = up.x; p.camUpDir[1] = up.y; p.camUpDir[2] = up.z;
    p.verticalFOV = cam.getVerticalFOV();
    p.numTriangles = scene->getTriangleCount();
    p.numSpheres = scene->getSphereCount();
    p.numMaterials = scene->getMaterialCount();
    p.numLights = 20;  // the reference passes its scratch size here; no kernel reads it
    p.maxRayTraceDepth = scene->getMaxRaytraceDepth();
    p.randomState = randomState;
    return p;
}

void Raytracer::doIteration(float frameTime) {
    (void)frameTime;
    rt_context *ctx = context.get();
    const auto t_frame = clock_t_::now();
    auto t0 = clock_t_::now();
    scene->updateScene();  // re-upload + device BVH build (S1) when the scene changed
    last.updateSceneMs = ms_since(t0);
    rt_scene_info info{};
    context.check(rt_scene_info_get(ctx, scene->getSceneId(), &info), "rt_scene_info_get");
    last.buildMs = info.build_ms;

    // a fresh random state per frame, from a seeded generator (the reference seeds std::mt19937 with the clock)
    std::mt19937 gen(randomState + iteration);
    const rt_params p = [&] {
        rt_params q = params();
        q.randomState = gen();
        return q;
    }();
    lastParams_ = p;
    const uint32_t spp = scene->getRaysPerPixel();
    const uint32_t npix = width * height;
    void *e0, *e1, *e2;
    context.check(rt_util_event_create(&e0), "event");
    context.check(rt_util_event_create(&e1), "event");
    context.check(rt_util_event_create(&e2), "event");
    context.check(rt_util_memset(segments, 0, 8), "memset");
    rt_util_event_record(e0, stream);
    context.check(rt_accum_clear(ctx, static_cast<float *>(accum), npix, stream), "rt_accum_clear");
    context.check(rt_render(ctx, scene->getSceneId(), &p, width, height, spp, 0, static_cast<float *>(accum),
                            static_cast<unsigned long long *>(segments), nullptr, stream),
                  "rt_render");
    rt_util_event_record(e1, stream);
    context.check(rt_resolve(ctx, static_cast<const float *>(accum), npix, spp, static_cast<float *>(resolved), stream),
                  "rt_resolve");
    rt_util_event_record(e2, stream);
    float tr = 0.0f, ts = 0.0f;
    context.check(rt_util_event_elapsed_ms(e0, e1, &tr), "event");
    context.check(rt_util_event_elapsed_ms(e1, e2, &ts), "event");
    rt_util_event_destroy(e0);
    rt_util_event_destroy(e1);
    rt_util_event_destroy(e2);
    last.traceMs = tr;
    last.resolveMs = ts;
    image_.resize((size_t)npix * 4);
    context.check(rt_util_memcpy_dtoh(image_.data(), resolved, image_.size() * sizeof(float)), "copy image");
    context.check(rt_util_memcpy_dtoh(&last.segments, segments, 8), "copy counter");
    last.frameMs = ms_since(t_frame);
    if (printTimings)
        std::printf("TIMINGS:\n\tupdateSceneTime: %.3f ms (device BVH build %.3f ms), traceTime: %.3f ms "
                    "(%llu segments, %.1f Msegments/s), resolveTime: %.3f ms, frameTime: %.3f ms\n",
                    last.updateSceneMs, last.buildMs, last.traceMs, last.segments,
                    last.traceMs > 0 ? last.segments / (last.traceMs * 1e3) : 0.0, last.resolveMs, last.frameMs);
    ++iteration;
}

void Raytracer::mainLoop(uint32_t frames) {
    auto current = clock_t_::now();
    for (uint32_t f = 0; f < frames; ++f) {
        auto now = clock_t_::now();
        double frameUs = std::chrono::duration<double, std::micro>(now - current).count();
        current = now;
        if (printTimings)
            std::printf("Frame Time(us): %.0f RaysPerPixel: %u Depth: %u\n", frameUs, scene->getRaysPerPixel(),
                        scene->getMaxRaytraceDepth());
        doIteration((float)frameUs);
    }
}

size_t Raytracer::sppSweep(const SppSweep &c, const std::string &csvPath) {
    if (c.runsBeforeIncrease == 0) throw std::invalid_argument("sppSweep: runsBeforeIncrease must be positive");
    if (c.increaseAmount == 0 && c.startRaysPerPixel <= c.maxRaysPerPixel)
        throw std::invalid_argument("sppSweep: increaseAmount must be positive");
    scene->setRaysPerPixel(c.startRaysPerPixel);
    std::vector<std::vector<long long>> times;  // microseconds, per level
    auto current = clock_t_::now();
    for (uint32_t it = 0;; ++it) {
        const auto now = clock_t_::now();
        const long long frameUs = std::chrono::duration_cast<std::chrono::microseconds>(now - current).count();
        current = now;
        if (printTimings)
            std::printf("Frame Time(us): %lld RaysPerPixel: %u Depth: %u\n", frameUs, scene->getRaysPerPixel(),
                        scene->getMaxRaytraceDepth());
        doIteration((float)frameUs);
        if (it == 0) continue;
        const size_t index = (it - 1) / c.runsBeforeIncrease;
        if (scene->getRaysPerPixel() > c.maxRaysPerPixel) break;
        if (times.size() == index) times.push_back({frameUs});
        else times[index].push_back(frameUs);
        if (it % c.runsBeforeIncrease == 0) scene->setRaysPerPixel(scene->getRaysPerPixel() + c.increaseAmount);
    }
    std::ofstream out(csvPath, std::ios::out | std::ios::trunc);
    if (!out) throw std::runtime_error("cannot write " + csvPath);
    size_t levels = 0;
    for (size_t i = 0; i < times.size(); ++i) {
        if (times[i].size() < c.runsBeforeIncrease) break;  // an unfinished level
        long long sum = 0;
        for (uint32_t j = 0; j < c.runsBeforeIncrease; ++j) sum += times[i][j];
        out << (i + 1) << ", " << sum / (long long)c.runsBeforeIncrease << "us,\n";
        ++levels;
    }
    return levels;
}

std::vector<float> Raytracer::accumulation() const {
    std::vector<float> out((size_t)width * height * 4);
    const_cast<rtgpu::Context &>(context).check(rt_util_memcpy_dtoh(out.data(), accum, out.size() * sizeof(float)),
                                                 "copy accumulation");
    return out;
}

void Raytracer::writePPM(const std::string &path) const {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    f << "P6\n" << width << " " << height << "\n255\n";
    std::vector<unsigned char> row((size_t)width * 3);
    for (uint32_t y = 0; y < height; ++y) {
        for (uint32_t x = 0; x < width; ++x)
            for (int c = 0; c < 3; ++c) {
                float v = image_.empty() ? 0.0f : image_[4 * ((size_t)y * width + x) + c];
                v = std::isnan(v) ? 0.0f : std::min(1.0f, std::max(0.0f, v));
                row[3 * x + c] = (unsigned char)(v * 255.0f + 0.5f);
            }
        f.write(reinterpret_cast<const char *>(row.data()), (std::streamsize)row.size());
    }
}

}  // namespace RaytracerBVHRenderer
