// api.hip -- the extern "C" boundary (include/rtgpu.h).  Exceptions never cross it.
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <map>
#include <vector>

#include "rt_internal.hpp"

using namespace rtgpu;

// byte layouts of the reference's SSBOs / UBO (RaytracerGPU_MastersProject/shaders/include/definitions.glsl,
// VulkanWrapper/SceneTypes.hpp, RaytracerBVH.hpp:31-42)
static_assert(sizeof(rt_triangle) == 64 && offsetof(rt_triangle, v1) == 16 && offsetof(rt_triangle, v2) == 32 &&
                  offsetof(rt_triangle, materialIndex) == 48 && offsetof(rt_triangle, modelIndex) == 52,
              "Triangle layout");
static_assert(sizeof(rt_sphere) == 32 && offsetof(rt_sphere, radius) == 16 && offsetof(rt_sphere, materialIndex) == 20 &&
                  offsetof(rt_sphere, modelIndex) == 24,
              "Sphere layout");
static_assert(sizeof(rt_material) == 32 && offsetof(rt_material, materialType) == 16, "Material layout");
static_assert(sizeof(rt_bvh_node) == 40 && offsetof(rt_bvh_node, left) == 24 && offsetof(rt_bvh_node, right) == 28 &&
                  offsetof(rt_bvh_node, primitiveIndex) == 32 && offsetof(rt_bvh_node, primitiveType) == 36,
              "HLBVHNode layout");
static_assert(sizeof(rt_params) == 80 && offsetof(rt_params, verticalFOV) == 48 &&
                  offsetof(rt_params, numTriangles) == 52 && offsetof(rt_params, numLights) == 64 &&
                  offsetof(rt_params, maxRayTraceDepth) == 68 && offsetof(rt_params, randomState) == 72,
              "ParameterUBO layout");
static_assert(sizeof(rt_ray) == 32 && sizeof(rt_hit) == 16, "ray/hit records");
static_assert(sizeof(LeafRecord) == 64, "leaf record");

namespace {

int hip_status(hipError_t e) {
    if (e == hipSuccess) return RT_OK;
    return e == hipErrorOutOfMemory ? RT_ERR_NO_MEMORY : RT_ERR_HIP;
}

}  // namespace

extern "C" {

int rt_create(rt_context **out, int device) {
    if (!out) return RT_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return RT_ERR_HIP;
    rt_context *ctx = nullptr;
    try {
        ctx = new rt_context();
    } catch (...) {
        return RT_ERR_NO_MEMORY;
    }
    ctx->device = device;
    int st = guarded(ctx, [&] {
        RT_HIP(hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, device));
        ctx->counters.alloc(16ull * rt_context::kCounterSlots);
        RT_HIP(hipMemset(ctx->counters.p, 0, ctx->counters.bytes));
        ctx->noalloc_scratch = static_cast<char *>(noalloc_scratch_address());  // this device's copy
    });
    if (st != RT_OK) {
        delete ctx;
        return st;
    }
    *out = ctx;
    return RT_OK;
}

void rt_destroy(rt_context *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    delete ctx;
}

const char *rt_last_error(const rt_context *ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int rt_scene_upload(rt_context *ctx, uint32_t *scene_id, const rt_triangle *tris, uint32_t ntri,
                    const rt_sphere *spheres, uint32_t nsph, const rt_material *materials, uint32_t nmat,
                    const float *models, uint32_t nmodels, const rt_bvh_node *nodes, uint32_t nnodes,
                    const rt_build_options *opts) {
    if (!ctx || !scene_id) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        rt_build_options o{};
        if (opts) o = *opts;
        auto sc = std::make_unique<Scene>();
        build_scene(ctx, *sc, tris, ntri, spheres, nsph, materials, nmat, models, nmodels, nodes, nnodes, o);
        uint32_t id = ctx->next_scene++;
        ctx->scenes[id] = std::move(sc);
        *scene_id = id;
    });
}

int rt_scene_release(rt_context *ctx, uint32_t scene_id) {
    if (!ctx) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        auto it = ctx->scenes.find(scene_id);
        if (it == ctx->scenes.end()) throw Error(RT_ERR_NOT_FOUND, "unknown scene id " + std::to_string(scene_id));
        if (it->second->users)
            throw Error(RT_ERR_BUSY, "scene " + std::to_string(scene_id) + " is used by " +
                                         std::to_string(it->second->users) + " server request(s)");
        ctx->wait_unused(*it->second);  // work enqueued on any stream may still read the scene
        ctx->scenes.erase(it);
    });
}

int rt_scene_info_get(rt_context *ctx, uint32_t scene_id, rt_scene_info *info) {
    if (!ctx || !info) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        const Scene &s = ctx->scene(scene_id);
        std::memset(info, 0, sizeof(*info));
        info->ntri = s.ntri; info->nsph = s.nsph; info->nmat = s.nmat; info->nprim = s.nprim;
        info->width = s.width; info->leaf_size = s.leaf_size; info->nwide = s.nwide; info->wide_depth = s.wide_depth;
        info->stack_bound = s.stack_bound; info->flags = s.flags;
        info->device_bytes = s.device_bytes();
        info->build_ms = s.build_ms;
    });
}

int rt_scene_get_nodes(rt_context *ctx, uint32_t scene_id, rt_bvh_node *host_out, uint32_t capacity) {
    if (!ctx || !host_out) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        const Scene &s = ctx->scene(scene_id);
        RT_REQUIRE(s.binary.p != nullptr || s.nprim == 0, "scene was built without RT_BUILD_KEEP_BINARY");
        uint64_t n = s.nprim ? 2ull * s.nprim - 1 : 0;
        RT_REQUIRE(capacity >= n, "capacity too small");
        if (n) RT_HIP(hipMemcpy(host_out, s.binary.p, sizeof(rt_bvh_node) * n, hipMemcpyDeviceToHost));
    });
}

int rt_scene_get_slots(rt_context *ctx, uint32_t scene_id, uint32_t *host_out, uint32_t capacity) {
    if (!ctx || !host_out) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        const Scene &s = ctx->scene(scene_id);
        RT_REQUIRE(capacity >= s.nprim, "capacity too small");
        if (s.nprim) RT_HIP(hipMemcpy(host_out, s.slot_prim.p, sizeof(uint32_t) * s.nprim, hipMemcpyDeviceToHost));
    });
}

int rt_scene_get_wide(rt_context *ctx, uint32_t scene_id, void *host_out, size_t capacity_bytes) {
    if (!ctx || !host_out) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        const Scene &s = ctx->scene(scene_id);
        RT_REQUIRE(capacity_bytes >= s.wide.bytes, "capacity too small");
        if (s.wide.bytes) RT_HIP(hipMemcpy(host_out, s.wide.p, s.wide.bytes, hipMemcpyDeviceToHost));
    });
}

int rt_trace(rt_context *ctx, uint32_t scene_id, const rt_ray *rays, uint32_t n, rt_hit *hits, uint32_t *counts,
             uint32_t flags, const rt_launch_config *cfg, void *stream) {
    if (!ctx) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        RT_REQUIRE((flags & ~kTraceFlags) == 0, "unknown trace flags");
        Scene &s = ctx->scene(scene_id);
        LaunchCfg c = resolve_cfg(ctx, s, cfg, (flags & RT_TRACE_ANY) != 0);
        launch_trace(ctx, s, rays, n, hits, counts, flags, c, static_cast<hipStream_t>(stream));
        s.note_use(static_cast<hipStream_t>(stream), ctx->record_use(static_cast<hipStream_t>(stream)));
    });
}

int rt_accum_clear(rt_context *ctx, float *accum, uint32_t npix, void *stream) {
    if (!ctx) return RT_ERR_INVALID;
    return guarded(ctx, [&] { launch_accum_clear(accum, npix, static_cast<hipStream_t>(stream)); });
}

int rt_render(rt_context *ctx, uint32_t scene_id, const rt_params *params, uint32_t width, uint32_t height,
              uint32_t spp, uint32_t flags, float *accum, unsigned long long *segments, const rt_launch_config *cfg,
              void *stream) {
    if (!ctx || !params) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        RT_REQUIRE((flags & ~kRenderFlags) == 0, "unknown render flags");
        Scene &s = ctx->scene(scene_id);
        LaunchCfg c = resolve_cfg(ctx, s, cfg, false);
        launch_render(ctx, s, *params, width, height, spp, flags, accum, segments, c,
                      static_cast<hipStream_t>(stream));
        s.note_use(static_cast<hipStream_t>(stream), ctx->record_use(static_cast<hipStream_t>(stream)));
    });
}

int rt_render_noalloc(rt_context *ctx, uint32_t scene_id, const rt_params *params, uint32_t width, uint32_t height,
                      uint32_t spp, uint32_t flags, float *accum, unsigned long long *segments,
                      const rt_launch_config *cfg, void *stream) {
    if (!ctx || !params) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        RT_REQUIRE((flags & ~kRenderFlags) == 0, "unknown render flags");
        RT_REQUIRE(ctx->noalloc_scratch != nullptr, "the static scratch array has no device address");
        Scene &s = ctx->scene(scene_id);
        LaunchCfg c = resolve_cfg(ctx, s, cfg, false);
        launch_render(ctx, s, *params, width, height, spp, flags, accum, segments, c, static_cast<hipStream_t>(stream),
                      ctx->noalloc_scratch, noalloc_scratch_bytes());
        s.note_use(static_cast<hipStream_t>(stream), ctx->record_use(static_cast<hipStream_t>(stream), false));
    });
}

int rt_resolve(rt_context *ctx, const float *accum, uint32_t npix, uint32_t spp, float *out_rgba, void *stream) {
    if (!ctx) return RT_ERR_INVALID;
    return guarded(ctx, [&] { launch_resolve(accum, npix, spp, out_rgba, static_cast<hipStream_t>(stream)); });
}

int rt_render_batch(rt_context *ctx, const rt_render_job *jobs, uint32_t njobs, unsigned long long *segments,
                    const rt_launch_config *cfg, void *stream) {
    if (!ctx || (njobs && !jobs)) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        if (njobs == 0) return;
        const LaunchCfg c = resolve_cfg(ctx, ctx->scene(jobs[0].scene_id), cfg, false);
        hipStream_t st = static_cast<hipStream_t>(stream);
        std::map<uint32_t, std::vector<RenderJob>> by_width;
        std::vector<std::pair<const Scene *, RenderJob>> waves;  // images rendered bounce by bounce, as rt_render would
        std::vector<std::pair<uintptr_t, uintptr_t>> ranges;
        std::vector<std::pair<RenderJob *, uint32_t>> split;  // (job, samples per item)
        uint64_t split_bytes = 0;
        for (uint32_t i = 0; i < njobs; ++i) {
            const rt_render_job &j = jobs[i];
            RT_REQUIRE((j.flags & ~kRenderFlags) == 0, "unknown render flags");
            const Scene &s = ctx->scene(j.scene_id);
            check_stack(s, c);
            RenderJob rj = make_render_job(s, j.params, j.width, j.height, j.spp, j.flags, j.accum);
            const uintptr_t b = reinterpret_cast<uintptr_t>(j.accum);
            ranges.emplace_back(b, b + 16ull * j.width * j.height);
            if (j.spp && wave_schedule(j.width, j.height, j.spp, j.flags)) waves.emplace_back(&s, rj);
            else by_width[s.width].push_back(rj);
        }
        std::sort(ranges.begin(), ranges.end());
        for (size_t i = 1; i < ranges.size(); ++i)
            RT_REQUIRE(ranges[i].first >= ranges[i - 1].second, "rt_render_batch: accumulation images overlap");
        for (auto &kv : by_width)
            for (RenderJob &rj : kv.second)
                if (const uint32_t g = split_group(ctx, rj.width, rj.height, rj.spp, rj.flags)) {
                    split.emplace_back(&rj, g);
                    split_bytes += 16ull * rj.width * rj.height * rj.spp;
                }
        std::unique_ptr<ScratchLease> lease;  // one block for every split job's samples
        if (split_bytes) lease = std::make_unique<ScratchLease>(ctx->scratch, split_bytes, st);
        char *scratch = lease ? lease->ptr() : nullptr;
        uint64_t off = 0;
        for (auto &[rj, g] : split) {
            make_split(*rj, reinterpret_cast<float4 *>(scratch + off), g);
            off += 16ull * rj->width * rj->height * rj->spp;
        }
        for (auto &kv : by_width) launch_render_jobs(ctx, kv.second, kv.first, segments, c, st);
        for (auto &[s, rj] : waves) launch_render_wave(ctx, *s, rj, segments, c, st, nullptr, 0);
        const int used = ctx->record_use(st);
        for (uint32_t i = 0; i < njobs; ++i) ctx->scene(jobs[i].scene_id).note_use(st, used);
    });
}

int rt_trace_batch(rt_context *ctx, const rt_batch_item *items, uint32_t nitems, uint32_t flags,
                   const rt_launch_config *cfg, void *stream) {
    if (!ctx) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        RT_REQUIRE((flags & ~kTraceFlags) == 0, "unknown trace flags");
        LaunchCfg c;
        const bool any = (flags & RT_TRACE_ANY) != 0;
        if (nitems) c = resolve_cfg(ctx, ctx->scene(items[0].scene_id), cfg, any);
        if (!cfg && ctx->has_batch_tuned) c = any ? ctx->batch_any : ctx->batch_closest;
        launch_trace_batch(ctx, items, nitems, flags, c, static_cast<hipStream_t>(stream));
        const int used = ctx->record_use(static_cast<hipStream_t>(stream));
        for (uint32_t i = 0; i < nitems; ++i) ctx->scene(items[i].scene_id).note_use(static_cast<hipStream_t>(stream), used);
    });
}

// ---------------- plumbing ----------------
int rt_util_malloc(void **ptr, size_t bytes) { return ptr ? hip_status(hipMalloc(ptr, bytes)) : RT_ERR_INVALID; }
int rt_util_free(void *ptr) { return hip_status(hipFree(ptr)); }
int rt_util_memcpy_htod(void *dst, const void *src, size_t bytes) {
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}
int rt_util_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}
int rt_util_memcpy_dtod(void *dst, const void *src, size_t bytes) {
    return hip_status(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
}
int rt_util_memset(void *ptr, int value, size_t bytes) { return hip_status(hipMemset(ptr, value, bytes)); }
int rt_util_synchronize(void) { return hip_status(hipDeviceSynchronize()); }
int rt_util_stream_create(void **stream) {
    return stream ? hip_status(hipStreamCreate(reinterpret_cast<hipStream_t *>(stream))) : RT_ERR_INVALID;
}
int rt_util_stream_destroy(void *stream) { return hip_status(hipStreamDestroy(static_cast<hipStream_t>(stream))); }
int rt_util_stream_synchronize(void *stream) {
    return hip_status(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
}
int rt_util_event_create(void **event) {
    return event ? hip_status(hipEventCreate(reinterpret_cast<hipEvent_t *>(event))) : RT_ERR_INVALID;
}
int rt_util_event_destroy(void *event) { return hip_status(hipEventDestroy(static_cast<hipEvent_t>(event))); }
int rt_util_event_record(void *event, void *stream) {
    return hip_status(hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream)));
}
int rt_util_event_elapsed_ms(void *start, void *end, float *ms) {
    if (!ms) return RT_ERR_INVALID;
    hipError_t e = hipEventSynchronize(static_cast<hipEvent_t>(end));
    if (e != hipSuccess) return hip_status(e);
    return hip_status(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end)));
}
int rt_util_device_count(int *count) { return count ? hip_status(hipGetDeviceCount(count)) : RT_ERR_INVALID; }
int rt_util_device_name(int device, char *buf, size_t len) {
    if (!buf || len == 0) return RT_ERR_INVALID;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return hip_status(e);
    char name[256] = {0};
    if (hipDeviceGetName(name, sizeof name, device) != hipSuccess || !name[0]) std::snprintf(name, sizeof name, "%s", prop.name);
    std::snprintf(buf, len, "%s (%s, %d CUs)", name[0] ? name : "AMD GPU", prop.gcnArchName, prop.multiProcessorCount);
    return RT_OK;
}

}  // extern "C"
