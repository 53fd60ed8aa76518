// tune.hip -- the autotuners: launch shape for one scene (rt_autotune), launch shape of
// the fused multi-scene launch (rt_autotune_batch), and BVH build parameters (node
// width, leaf size) for a set of sample scenes (rt_autotune_build).
//
// Every candidate is timed with HIP events on a private stream: one warm-up launch,
// then the median of three.
#include <algorithm>
#include <functional>
#include <memory>
#include <vector>

#include "rt_internal.hpp"
#include "traverse.hpp"

namespace rtgpu {
namespace {

class Timer {
    hipStream_t stream_ = nullptr;
    hipEvent_t e0_ = nullptr, e1_ = nullptr;

public:
    Timer() {
        RT_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        RT_HIP(hipEventCreate(&e0_));
        RT_HIP(hipEventCreate(&e1_));
    }
    ~Timer() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        if (e0_) (void)hipEventDestroy(e0_);
        if (e1_) (void)hipEventDestroy(e1_);
        if (stream_) (void)hipStreamDestroy(stream_);
    }
    Timer(const Timer &) = delete;
    Timer &operator=(const Timer &) = delete;
    hipStream_t stream() const { return stream_; }
    // median of three timed runs after one warm-up
    float median_ms(const std::function<void(hipStream_t)> &launch) {
        launch(stream_);
        float t[3];
        for (float &ms : t) {
            RT_HIP(hipEventRecord(e0_, stream_));
            launch(stream_);
            RT_HIP(hipEventRecord(e1_, stream_));
            RT_HIP(hipEventSynchronize(e1_));
            RT_HIP(hipEventElapsedTime(&ms, e0_, e1_));
        }
        std::sort(t, t + 3);
        return t[1];
    }
};

// Launch shapes whose stack fits every tree of depth bound `stack_bound`.
std::vector<LaunchCfg> launch_space(uint32_t stack_bound, bool with_direct, bool full) {
    std::vector<LaunchCfg> v;
    const std::vector<uint32_t> stacks = full ? std::vector<uint32_t>{4u, 8u, 16u} : std::vector<uint32_t>{8u, 16u};
    const std::vector<uint32_t> refills = full ? std::vector<uint32_t>{8u, 16u, 32u} : std::vector<uint32_t>{16u};
    for (uint32_t block : {64u, 128u, 256u})
        for (uint32_t st : stacks) {
            if (stack_bound > st + kSpillCap) continue;
            if (with_direct) v.push_back(LaunchCfg{block, st, 0u, 16u});
            for (uint32_t refill : refills) v.push_back(LaunchCfg{block, st, 1u, refill});
        }
    RT_REQUIRE(!v.empty(), "no launch configuration fits the trees' stack bound");
    return v;
}

struct Best {
    LaunchCfg cfg;
    float ms = 1e30f;
    uint32_t tried = 0;
};

Best search(Timer &timer, const std::vector<LaunchCfg> &space,
            const std::function<void(const LaunchCfg &, hipStream_t)> &launch) {
    Best b;
    b.cfg = space[0];
    for (const LaunchCfg &c : space) {
        const float ms = timer.median_ms([&](hipStream_t s) { launch(c, s); });
        b.tried++;
        if (ms < b.ms) {
            b.ms = ms;
            b.cfg = c;
        }
    }
    return b;
}

rt_launch_config to_abi(const LaunchCfg &c) { return rt_launch_config{c.block, c.lds_stack, c.persistent, c.refill}; }

}  // namespace
}  // namespace rtgpu

using namespace rtgpu;

extern "C" {

int rt_autotune(rt_context *ctx, uint32_t scene_id, const rt_ray *rays, uint32_t n, rt_hit *hits, uint32_t flags,
                rt_tune_result *out) {
    if (!ctx) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        RT_REQUIRE(n > 0, "autotune needs rays");
        RT_REQUIRE((flags & ~kTraceFlags) == 0, "unknown trace flags");
        Scene &s = ctx->scene(scene_id);
        const bool any = (flags & RT_TRACE_ANY) != 0;
        Timer timer;
        const Best b = search(timer, launch_space(s.stack_bound, true, true), [&](const LaunchCfg &c, hipStream_t st) {
            launch_trace(ctx, s, rays, n, hits, nullptr, flags, c, st);
        });
        if (!s.has_tuned) {
            s.tuned_closest = LaunchCfg{};
            s.tuned_any = LaunchCfg{};
        }
        (any ? s.tuned_any : s.tuned_closest) = b.cfg;
        s.has_tuned = true;
        if (out) {
            out->cfg = to_abi(b.cfg);
            out->ms = b.ms;
            out->rays_per_s = b.ms > 0.0f ? (float)(n / (b.ms * 1e-3)) : 0.0f;
            out->tried = b.tried;
        }
    });
}

int rt_autotune_batch(rt_context *ctx, const rt_batch_item *items, uint32_t nitems, uint32_t flags,
                      rt_tune_result *out) {
    if (!ctx || (nitems && !items)) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        RT_REQUIRE((flags & ~kTraceFlags) == 0, "unknown trace flags");
        std::vector<TraceItem> v;
        uint32_t bound = 0;
        uint64_t total = 0;
        for (uint32_t i = 0; i < nitems; ++i) {
            const Scene &s = ctx->scene(items[i].scene_id);
            v.push_back(TraceItem{&s, items[i].rays, items[i].hits, items[i].nrays});
            bound = std::max(bound, s.stack_bound);
            total += items[i].nrays;
        }
        RT_REQUIRE(total > 0, "autotune needs rays");
        Timer timer;
        const Best b = search(timer, launch_space(bound, false, true), [&](const LaunchCfg &c, hipStream_t st) {
            launch_trace_items(ctx, v, flags, c, st);
        });
        const bool any = (flags & RT_TRACE_ANY) != 0;
        if (!ctx->has_batch_tuned) ctx->batch_closest = ctx->batch_any = LaunchCfg{};
        (any ? ctx->batch_any : ctx->batch_closest) = b.cfg;
        ctx->has_batch_tuned = true;
        if (out) {
            out->cfg = to_abi(b.cfg);
            out->ms = b.ms;
            out->rays_per_s = b.ms > 0.0f ? (float)(total / (b.ms * 1e-3)) : 0.0f;
            out->tried = b.tried;
        }
    });
}

int rt_autotune_build(rt_context *ctx, const rt_tune_scene *scenes, uint32_t nscenes, uint32_t flags,
                      rt_build_candidate *results, uint32_t capacity, uint32_t *ncandidates, rt_build_options *best) {
    if (!ctx || (nscenes && !scenes)) return RT_ERR_INVALID;
    return guarded(ctx, [&] {
        RT_REQUIRE(nscenes > 0, "autotune needs sample scenes");
        RT_REQUIRE((flags & ~kTraceFlags) == 0, "unknown trace flags");
        uint64_t total = 0;
        for (uint32_t i = 0; i < nscenes; ++i) {
            RT_REQUIRE(scenes[i].nrays == 0 || (scenes[i].rays && scenes[i].hits), "sample scene without device rays");
            total += scenes[i].nrays;
        }
        RT_REQUIRE(total > 0, "autotune needs rays");
        Timer timer;
        std::vector<rt_build_candidate> cands;
        for (uint32_t width : {2u, 4u, 8u})
            for (uint32_t leaf : {1u, 2u, 4u}) {
                const rt_build_options opts{0u, width, leaf, 0u};
                std::vector<std::unique_ptr<Scene>> built;
                std::vector<TraceItem> items;
                float build_ms = 0.0f;
                uint64_t bytes = 0;
                uint32_t bound = 0;
                for (uint32_t i = 0; i < nscenes; ++i) {
                    const rt_tune_scene &t = scenes[i];
                    auto sc = std::make_unique<Scene>();
                    build_scene(ctx, *sc, t.tris, t.ntri, t.spheres, t.nsph, t.materials, t.nmat, nullptr, 0, nullptr,
                                0, opts);
                    build_ms += sc->build_ms;
                    bytes += sc->device_bytes();
                    bound = std::max(bound, sc->stack_bound);
                    items.push_back(TraceItem{sc.get(), t.rays, t.hits, t.nrays});
                    built.push_back(std::move(sc));
                }
                const Best b = search(timer, launch_space(bound, false, false), [&](const LaunchCfg &c, hipStream_t st) {
                    launch_trace_items(ctx, items, flags, c, st);
                });
                RT_HIP(hipStreamSynchronize(timer.stream()));  // the trees are freed next
                cands.push_back(rt_build_candidate{width, leaf, build_ms, b.ms, to_abi(b.cfg), bytes});
            }
        std::stable_sort(cands.begin(), cands.end(),
                         [](const rt_build_candidate &a, const rt_build_candidate &b) { return a.trace_ms < b.trace_ms; });
        if (results)
            for (uint32_t i = 0; i < std::min<uint32_t>(capacity, (uint32_t)cands.size()); ++i) results[i] = cands[i];
        if (ncandidates) *ncandidates = (uint32_t)cands.size();
        if (best) *best = rt_build_options{0u, cands[0].width, cands[0].leaf_size, 0u};
    });
}

}  // extern "C"
