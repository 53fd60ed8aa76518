// server.hip -- the serving loop: resident scenes, budgeted admission, continuous batching.
//
// Requests wait in a FIFO queue.  The head is admitted when a slot is free and the
// request's buffers fit in the first half of the server's device arena (first-fit
// allocation in one hipMalloc made at creation, so admission never calls the allocator).
// Admitted requests advance in scheduling steps on two lanes, each a stream with up to
// `pipeline` steps enqueued ahead of the device:
//   - the big lane carries renders of at least kWaveMinSamples pixel-samples (rt_render's
//     threshold for filling the device), the oldest alone, bounce by bounce as rt_render runs
//     it, in slices of the step's budget of device time (rates measured from earlier steps);
//   - the small lane carries the other renders, together in one persistent launch
//     (launch_render_jobs, split layout), whole where they fit the budget, plus one fused
//     ray-query launch over slices of every admitted query (launch_trace_items).
// The lanes' kernels share the device: a small step's kernels take compute units as the big
// lane's bounce kernels release them.  A big render's last bounces run in one persistent
// launch; while small requests keep arriving its slices are single yieldable passes, and
// enqueuing a small step asks that launch to give the device up once the pass has run
// yield_min_ms (its waves finish the paths they hold and exit).  The pass is resumed on the big
// lane after the small step, so a small request waits for the drain of the paths in flight, not
// for a whole slice.  A request keeps its lane for life, so its steps stay in order on one stream.
// A step's path state or per-sample storage comes from the second half of the arena for the
// step's lifetime.
// A request's last slice is followed by its resolve; when that step's event completes
// the result is copied out and its arena range freed.
//
// A render's samples may be cut into any number of slices without changing its image
// (each pixel's chain runs through its accumulator), and a ray's result does not
// depend on the launch it is traced in, so results are the same alone or in a mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "rt_internal.hpp"

namespace rtgpu {
namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over [0, size) with 256-byte granularity.
class Arena {
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    uint64_t size_ = 0, used_ = 0, peak_ = 0;

public:
    static constexpr uint64_t kAlign = 256;
    static uint64_t round(uint64_t n) { return (n + kAlign - 1) / kAlign * kAlign; }
    void reset(uint64_t size) {
        size_ = size / kAlign * kAlign;
        free_.clear();
        if (size_) free_[0] = size_;
        used_ = peak_ = 0;
    }
    bool alloc(uint64_t n, uint64_t &off) {
        n = round(std::max<uint64_t>(n, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const uint64_t rest = it->second - n;
            free_.erase(it);
            if (rest) free_[off + n] = rest;
            used_ += n;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        n = round(std::max<uint64_t>(n, 1));
        used_ -= n;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + n == next->first) {  // merge with the following range
            n += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {  // merge with the preceding range
                prev->second += n;
                return;
            }
        }
        free_[off] = n;
    }
    uint64_t size() const { return size_; }
    uint64_t peak() const { return peak_; }
};

enum class State { Queued, Active, Finishing, Done };

struct Request {
    uint64_t id = 0;
    rt_request req{};
    State state = State::Queued;
    Scene *scene = nullptr;
    uint64_t bytes = 0, offset = 0;
    uint32_t remaining = 0;  // samples (render) or rays (query) not yet given to a step
    uint32_t done_units = 0;
    bool big_lane = false;   // its steps run on the device-filling renders' stream (set at admission)
    bool primary_ready = false;  // a big-lane render's primary hits have been traced (see request_bytes)
    rt_request_stats st{};
};

struct Step {
    hipEvent_t start = nullptr, end = nullptr;
    std::shared_ptr<WaveResume> resume;  // a yieldable pass: resumed at retirement if it yielded
    uint32_t *status = nullptr;          // host copy of that pass's counters, written before `end`
    std::vector<std::shared_ptr<Request>> finishing;
    std::vector<std::shared_ptr<Request>> members;
    uint64_t samples = 0, rays = 0;
    uint64_t scratch_off = 0, scratch_bytes = 0;  // per-sample storage of the step's renders
    bool big_lane = false;    // enqueued on the device-filling renders' stream
    bool overlapped = false;  // shared the device with a step of the other stream
};

constexpr uint32_t kStatusSlots = 64;  // status slots of yieldable passes (at most two are ever in use)

}  // namespace
}  // namespace rtgpu

struct rt_server {
    rt_context *ctx = nullptr;
    rt_server_config cfg{};
    rtgpu::DevBuf arena_mem;
    rtgpu::Arena arena;    // first half of arena_mem: admitted requests' images, rays and hits
    rtgpu::Arena scratch;  // second half: per-sample storage of the steps in flight
    uint64_t scratch_base = 0;
    rtgpu::DevBuf segments;  // u64 device counter of traced render segments
    // Fine-grained host memory: [0] counts yield requests (the big lane's yieldable finish launches
    // read it), then kStatusSlots slots of 4 words that receive yieldable passes' counters.
    uint32_t *pinned = nullptr;
    uint32_t *pinned_dev = nullptr;  // its device address
    uint32_t yield_gen = 0;          // the value in pinned[0]
    uint32_t status_next = 0;
    bool yield_pending = false;      // a small step was enqueued while a yieldable pass was in flight
    double yield_min_ms = 0.5;       // a pass runs this long before a small step may make it yield
    // Two lanes: device-filling renders run on `stream`, everything else (small renders, ray
    // queries) on `small_stream`, so a small request never waits behind a device-filling step:
    // its kernels take compute units as the other lane's kernels release them.  The unbatched
    // mode uses `stream` only.
    hipStream_t stream = nullptr, small_stream = nullptr;
    hipEvent_t t0_event = nullptr;  // recorded at creation: the origin of step intervals
    std::deque<std::pair<float, float>> recent;  // device intervals of the last retired steps (busy_ms)
    std::vector<hipEvent_t> event_pool;

    std::mutex m;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<rtgpu::Request>> queue;
    std::vector<std::shared_ptr<rtgpu::Request>> active;
    std::deque<rtgpu::Step> inflight;
    std::unordered_map<uint64_t, std::shared_ptr<rtgpu::Request>> all;
    uint64_t next_id = 1;
    bool stop = false;
    int fatal = RT_OK;
    std::string last_error;

    // rates measured from retired steps (work units per microsecond of device time)
    double sample_rate = 1000.0, ray_rate = 2000.0;
    // host time (ms since creation) of the last admission of a request that is not a device-filling
    // render: while such requests keep arriving, a device-filling render advances one yieldable
    // pass at a time (or, with RT_SERVER_NO_YIELD, in short slices)
    double last_small_ms = -1e30;
    double last_big_ms = -1e30;  // host time of the last big-lane step enqueued
    double slice_div = 8.0;      // RT_SERVER_NO_YIELD: big-lane slices while small requests arrive, budget / slice_div
    double small_div = 5.0;      // a fused small-lane step's share of the budget: budget / small_div
    rt_server_stats stats{};
    rtgpu::Clock::time_point t0;
    std::thread worker;
    // Results bound for the callers' host buffers are copied on a thread of their own (completer_main),
    // so the worker keeps scheduling while a large image crosses the host link.
    std::deque<std::shared_ptr<rtgpu::Request>> completing;
    std::condition_variable cv_complete;
    bool completer_stop = false;
    std::thread completer;

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(rtgpu::Clock::now() - t0).count();
    }
};

namespace rtgpu {
namespace {

// A request's own arena range: a render's accumulation image, and for a device-filling render of
// the batched server its pixels' primary hits after it (16 B each, traced once for all its
// slices); a ray query's rays and hits.  Step storage comes from the arena's other half, and the
// unbatched server renders whole requests through rt_render's own scratch.
bool big_render(const rt_request &r, bool serial) {
    return !serial && r.kind == RT_REQ_RENDER && (uint64_t)r.width * r.height * r.spp >= kWaveMinSamples;
}

uint64_t request_bytes(const rt_request &r, bool serial) {
    if (r.kind != RT_REQ_RENDER)  // a query on device buffers keeps no copies in the arena
        return (r.trace_flags & RT_REQ_DEVICE_BUFFERS) ? 0 : Arena::round(32ull * r.nrays) + Arena::round(16ull * r.nrays);
    const uint64_t image = Arena::round(16ull * r.width * r.height);
    return big_render(r, serial) ? 2 * image : image;
}

// resident lanes of the device, about 1024 per CU
uint64_t device_lanes(const rt_server *s) { return 1024ull * (uint64_t)s->ctx->num_cus; }

hipEvent_t take_event(rt_server *s) {
    if (!s->event_pool.empty()) {
        hipEvent_t e = s->event_pool.back();
        s->event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    RT_HIP(hipEventCreate(&e));
    return e;
}

char *arena_ptr(rt_server *s, uint64_t off) { return static_cast<char *>(s->arena_mem.p) + off; }

uint32_t *take_status(rt_server *s) {
    static_assert(kWaveCounterWords <= 4, "a status slot holds 4 words");
    return s->pinned + 4 + 4 * (s->status_next++ % kStatusSlots);
}

bool yieldable_inflight(const rt_server *s) {
    for (const Step &st : s->inflight)
        if (st.big_lane && st.resume) return true;
    return false;
}

// Asks the big lane's yieldable finish launches enqueued so far to give the device up.
void request_yield(rt_server *s) {
    s->yield_gen++;
    __atomic_store_n(s->pinned, s->yield_gen, __ATOMIC_SEQ_CST);
}

// How long the in-flight yieldable pass has run on the device (0 before it starts; the device
// clock of its start event against the server's origin event, which is also the host clock's).
double yieldable_run_ms(rt_server *s) {
    for (const Step &st : s->inflight) {
        if (!st.big_lane || !st.resume) continue;
        float a = 0.0f;
        if (hipEventQuery(st.start) != hipSuccess || hipEventElapsedTime(&a, s->t0_event, st.start) != hipSuccess)
            return 0.0;
        return std::max(0.0, s->now_ms() - (double)a);
    }
    return 0.0;
}

// The big lane's next work starts after the small steps in flight (the small lane runs in order,
// so its last step's end covers them).  Called with the context lock held.
void after_small_steps(rt_server *s) {
    for (auto it = s->inflight.rbegin(); it != s->inflight.rend(); ++it)
        if (!it->big_lane) {
            RT_HIP(hipStreamWaitEvent(s->stream, it->end, 0));
            return;
        }
}

void resolve_finishing(rt_server *s, const std::vector<std::shared_ptr<Request>> &finishing, hipStream_t stream,
                       const uint32_t *gate) {
    for (auto &r : finishing) {
        if (r->req.kind != RT_REQ_RENDER || (r->req.render_flags & RT_REQ_OUT_ACCUM)) continue;
        float *accum = reinterpret_cast<float *>(arena_ptr(s, r->offset));
        launch_resolve(accum, r->req.width * r->req.height, r->req.spp, accum, stream, gate);  // in place
    }
}

hipStream_t lane_stream(const rt_server *s, bool big_lane) {
    return big_lane || (s->cfg.flags & RT_SERVER_SERIAL) ? s->stream : s->small_stream;
}

size_t lane_inflight(const rt_server *s, bool big_lane) {
    size_t n = 0;
    for (const Step &st : s->inflight) n += st.big_lane == big_lane ? 1 : 0;
    return n;
}

// Admits queue heads while a slot and arena space are free.  Called with s->m held.  A slot is held
// by a request with work not yet enqueued; one whose last slice is on the device holds only its arena
// range, so the next step can be formed while that one runs (otherwise a step that takes every
// admitted query whole leaves the lane idle from its end until the host has seen it finish).
void admit(rt_server *s) {
    const uint32_t slots = (s->cfg.flags & RT_SERVER_SERIAL) ? std::max(1u, s->cfg.pipeline) : s->cfg.max_active;
    size_t busy = 0;
    for (const auto &r : s->active) busy += (r->state == State::Active && r->remaining > 0) ? 1 : 0;
    while (!s->queue.empty() && busy < slots) {
        std::shared_ptr<Request> r = s->queue.front();
        uint64_t off = 0;
        if (!s->arena.alloc(r->bytes, off)) break;  // arrival order: the head waits for memory
        s->queue.pop_front();
        r->offset = off;
        r->state = State::Active;
        r->st.admit_ms = s->now_ms();
        // A request keeps its lane for life, so its steps run in order on one stream.
        r->big_lane = big_render(r->req, (s->cfg.flags & RT_SERVER_SERIAL) != 0);
        if (!r->big_lane) s->last_small_ms = r->st.admit_ms;
        {
            std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
            hipStream_t st = lane_stream(s, r->big_lane);
            if (r->req.kind == RT_REQ_RENDER) {
                launch_accum_clear(reinterpret_cast<float *>(arena_ptr(s, off)), r->req.width * r->req.height, st);
            } else if (r->req.nrays && !(r->req.trace_flags & RT_REQ_DEVICE_BUFFERS)) {
                RT_HIP(hipMemcpyAsync(arena_ptr(s, off), r->req.rays, 32ull * r->req.nrays, hipMemcpyHostToDevice, st));
            }
        }
        s->active.push_back(r);
        busy += r->remaining > 0 ? 1 : 0;
    }
}

// Builds and enqueues one step of a lane.  Called with s->m held; returns false when nothing was
// scheduled.  The unbatched mode has one lane (big_lane true) that takes every request.
bool enqueue_step(rt_server *s, bool big_lane) {
    const bool serial = (s->cfg.flags & RT_SERVER_SERIAL) != 0;
    std::vector<std::shared_ptr<Request>> renders, queries;
    for (auto &r : s->active) {
        if (r->state != State::Active || r->remaining == 0 || (!serial && r->big_lane != big_lane)) continue;
        (r->req.kind == RT_REQ_RENDER ? renders : queries).push_back(r);
    }
    if (renders.empty() && queries.empty()) return false;
    if (serial) {  // the oldest admitted request, whole
        std::shared_ptr<Request> first;
        for (auto &r : s->active)
            if (r->state == State::Active && r->remaining > 0) { first = r; break; }
        renders.clear();
        queries.clear();
        (first->req.kind == RT_REQ_RENDER ? renders : queries).push_back(first);
    }
    const hipStream_t stream = lane_stream(s, big_lane);
    // Renders this step, in admission order: whole while they fit the step's sample budget;
    // the first that does not fit gets the number of samples that does (at least one when it
    // is the step's first render).
    //
    // Small lane: the admitted small renders are fused into one persistent launch, every job in
    // the split layout with one group size for the launch, chosen so the launch has ~32 work
    // items per resident lane: its drain (lanes finishing their last items) stays short and each
    // item traces its first segment once for its group of samples.  Such a step finishes all its
    // members together, so it gets a fifth of the budget (RTGPU_SERVER_SMALL_DIV): its first
    // members are not held long by its last.  Ray queries ride along (slices of every admitted query).
    //
    // Big lane: the oldest device-filling render (at least kWaveMinSamples pixel-samples at
    // admission, rt_render's own threshold) runs alone, bounce by bounce as rt_render would run
    // it (below), in slices of the budget.  Its last bounces run in one persistent launch that
    // holds the compute units to its end, so while small requests keep arriving (one admitted
    // within the last two step budgets) its slices are single passes whose finish launch yields
    // to small steps (or, with RT_SERVER_NO_YIELD, slices of budget / slice_div), and the worker
    // keeps at most one of them in flight and none while a small step runs (worker_main).
    //
    // Storage (path state, or per-sample colours) comes from the scratch half of the arena and
    // returns there when the step retires.
    double budget = std::max(1.0, s->sample_rate * s->cfg.step_us);
    std::vector<std::pair<std::shared_ptr<Request>, uint32_t>> plan;
    uint64_t total = 0;
    auto npix_of = [](const Request &r) { return (uint64_t)r.req.width * r.req.height; };
    const bool small_recent = s->now_ms() - s->last_small_ms < 2.0 * s->cfg.step_us / 1000.0;
    const bool big_step = !serial && big_lane && !renders.empty();
    const bool yields = !serial && !(s->cfg.flags & RT_SERVER_NO_YIELD);
    if (big_step) renders.resize(1);
    else if (!serial && !renders.empty()) budget = std::max(1.0, budget / s->small_div);
    if (big_step && small_recent && !yields) budget = std::max(1.0, budget / s->slice_div);
    for (auto &r : renders) {
        const uint64_t npix = npix_of(*r);
        uint32_t q = r->remaining;
        if (!serial) {
            if (s->cfg.max_spp_step) q = std::min(q, s->cfg.max_spp_step);
            const double room = budget - (double)total;
            if ((double)(npix * q) > room) {
                const uint32_t fit = (uint32_t)std::max(0.0, std::floor(room / (double)npix));
                q = plan.empty() ? std::max(1u, std::min(q, fit)) : std::min(q, fit);
            }
            if (q == 0) break;
        }
        plan.emplace_back(r, q);
        total += npix * q;
        if (!serial && (double)total >= budget) break;
    }
    // A big render alone runs bounce by bounce, like rt_render would run it, unless its
    // flags ask for the persistent launch; its path storage comes from the scratch half.
    const uint32_t layout_flags = RT_RENDER_SPLIT_SAMPLES | RT_RENDER_NO_SPLIT | RT_RENDER_PERSISTENT_PATHS;
    bool wave = big_step && plan.size() == 1 && !(plan[0].first->req.render_flags & layout_flags);
    uint64_t scratch_off = 0, scratch_bytes = 0;
    if (wave) {  // halve the slice until its path storage fits; else the persistent launch below
        auto &[r, q] = plan[0];
        const uint64_t npix = npix_of(*r);
        // while small requests arrive, one pass per step: its finish launch can yield to them
        if (yields && small_recent) q = (uint32_t)std::min<uint64_t>(q, per_pass_limit(npix));
        auto need = [&](uint64_t samples) { return wave_scratch_bytes(npix * std::min(samples, per_pass_limit(npix))); };
        uint32_t qq = q;
        bool fits = s->scratch.alloc(need(qq), scratch_off);
        while (!fits && qq > 1) {
            qq /= 2;
            fits = s->scratch.alloc(need(qq), scratch_off);
        }
        if (fits) {
            q = qq;
            total = npix * q;
            scratch_bytes = need(q);
        } else {
            wave = false;
        }
    }
    if (!wave && !serial) {
        while (!plan.empty() && !s->scratch.alloc(16ull * total, scratch_off)) {  // shrink until it fits
            auto &[r, q] = plan.back();
            if (plan.size() > 1 || q == 1) {
                total -= npix_of(*r) * q;
                plan.pop_back();
            } else {
                total -= npix_of(*r) * (q - q / 2);
                q /= 2;
            }
        }
    }
    if (plan.empty() && queries.empty()) return false;  // storage is held by steps in flight
    Step step;
    step.start = take_event(s);
    step.end = take_event(s);
    if (!serial && !plan.empty()) {
        step.scratch_off = scratch_off;
        step.scratch_bytes = wave ? scratch_bytes : 16ull * total;
    }
    std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
    const bool yieldable = wave && yields && small_recent;
    RT_HIP(hipEventRecord(step.start, stream));
    const uint32_t group = (uint32_t)std::max<uint64_t>(1, total / (32ull * device_lanes(s)));
    char *storage = static_cast<char *>(s->arena_mem.p) + s->scratch_base + scratch_off;
    std::map<uint32_t, std::vector<RenderJob>> by_width;
    LaunchCfg rc;
    bool rc_set = false;
    for (auto &[r, q] : plan) {
        const uint64_t npix = npix_of(*r);
        float *accum = reinterpret_cast<float *>(arena_ptr(s, r->offset));
        const uint32_t kflags = r->req.render_flags & ~RT_REQ_OUT_ACCUM;
        if (!rc_set) {
            rc = resolve_cfg(s->ctx, *r->scene, &s->cfg.render_cfg, false);
            if (!s->cfg.render_cfg.persistent) rc.persistent = 1;
            rc_set = true;
        }
        check_stack(*r->scene, rc);
        if (serial) {
            launch_render(s->ctx, *r->scene, r->req.params, r->req.width, r->req.height, q, kflags, accum,
                          s->segments.as<unsigned long long>(), rc, stream);
        } else if (wave) {
            RenderJob j = make_render_job(*r->scene, r->req.params, r->req.width, r->req.height, q, kflags, accum);
            // the primary hits, kept after the image, are traced by the render's first bounce-by-bounce slice
            rt_hit *primary = reinterpret_cast<rt_hit *>(arena_ptr(s, r->offset + Arena::round(16ull * npix)));
            const WaveYield wy{s->pinned_dev, s->yield_gen + 1};  // yields to small steps enqueued from now on
            launch_render_wave(s->ctx, *r->scene, j, s->segments.as<unsigned long long>(), rc, stream, storage,
                               scratch_bytes, r->big_lane ? primary : nullptr, r->primary_ready,
                               yieldable ? &wy : nullptr, yieldable ? &step.resume : nullptr);
            r->primary_ready = r->big_lane;
        } else {
            RenderJob j = make_render_job(*r->scene, r->req.params, r->req.width, r->req.height, q, kflags, accum);
            make_split(j, reinterpret_cast<float4 *>(storage), std::min(group, q));
            storage += 16ull * npix * q;
            by_width[r->scene->width].push_back(j);
        }
        r->remaining -= q;
        r->done_units += q;
        r->st.steps++;
        step.samples += npix * q;
        step.members.push_back(r);
    }
    for (auto &kv : by_width)
        launch_render_jobs(s->ctx, kv.second, kv.first, s->segments.as<unsigned long long>(), rc, stream);
    if (!queries.empty()) {
        const double rshare = s->ray_rate * s->cfg.step_us / (double)queries.size();
        std::vector<TraceItem> groups[4];  // by trace flags: closest/any x refined/fp32
        for (auto &r : queries) {
            uint32_t q = serial ? r->remaining
                                : (uint32_t)std::min<double>(r->remaining, std::max(65536.0, std::floor(rshare)));
            q = std::max(1u, std::min(q, r->remaining));
            const uint32_t first = r->done_units;
            const bool dev = (r->req.trace_flags & RT_REQ_DEVICE_BUFFERS) != 0;
            const rt_ray *rays = (dev ? r->req.rays : reinterpret_cast<const rt_ray *>(arena_ptr(s, r->offset))) + first;
            rt_hit *hits = (dev ? static_cast<rt_hit *>(r->req.out)
                                : reinterpret_cast<rt_hit *>(arena_ptr(s, r->offset + Arena::round(32ull * r->req.nrays)))) +
                           first;
            groups[r->req.trace_flags & kTraceFlags].push_back(TraceItem{r->scene, rays, hits, q});
            r->remaining -= q;
            r->done_units += q;
            r->st.steps++;
            step.rays += q;
            step.members.push_back(r);
        }
        for (uint32_t flags = 0; flags < 4; ++flags) {
            auto &items = groups[flags];
            if (items.empty()) continue;
            LaunchCfg tc = resolve_cfg(s->ctx, *items[0].sc, &s->cfg.trace_cfg, (flags & RT_TRACE_ANY) != 0);
            if (!s->cfg.trace_cfg.persistent) tc.persistent = 1;
            // a lone query goes through the single-scene kernel: the fused launch's item table and
            // per-chunk item lookup only pay off over several
            if (serial || items.size() == 1)
                launch_trace(s->ctx, *items[0].sc, items[0].rays, items[0].n, items[0].hits, nullptr, flags, tc, stream);
            else launch_trace_items(s->ctx, items, flags, tc, stream);
        }
    }
    for (auto &r : step.members) {
        if (r->remaining) continue;
        r->state = State::Finishing;
        step.finishing.push_back(r);
    }
    // a yieldable pass's resolve runs only if the pass completed; otherwise its resume step's does
    resolve_finishing(s, step.finishing, stream, step.resume ? wave_counters(*step.resume) : nullptr);
    if (step.resume) {
        step.status = take_status(s);
        RT_HIP(hipMemcpyAsync(step.status, wave_counters(*step.resume), kWaveCounterWords * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, stream));
    }
    RT_HIP(hipEventRecord(step.end, stream));
    s->stats.steps++;
    step.big_lane = big_lane;
    for (Step &o : s->inflight)
        if (o.big_lane != step.big_lane) o.overlapped = step.overlapped = true;
    s->inflight.push_back(std::move(step));
    return true;
}

// Device time covered by [a, b] and not by the recently retired steps' intervals: the steps of the
// two lanes overlap, and busy_ms counts the time the device ran any step.
double uncovered_ms(const std::deque<std::pair<float, float>> &recent, float a, float b) {
    std::vector<std::pair<float, float>> cut;
    for (const auto &iv : recent)
        if (iv.second > a && iv.first < b) cut.emplace_back(std::max(iv.first, a), std::min(iv.second, b));
    std::sort(cut.begin(), cut.end());
    double covered = 0.0, upto = a;
    for (const auto &iv : cut) {
        if (iv.second <= upto) continue;
        covered += iv.second - std::max<double>(iv.first, upto);
        upto = iv.second;
    }
    return std::max(0.0, (double)(b - a) - covered);
}

// A yielded pass goes on in a step of its own on the big lane, after the small steps in flight; the
// requests it finishes and its path storage move to that step.  Called with s->m held.
void resume_step(rt_server *s, Step &from) {
    Step rs;
    rs.start = take_event(s);
    rs.end = take_event(s);
    rs.resume = from.resume;
    rs.finishing = std::move(from.finishing);
    from.finishing.clear();
    rs.scratch_off = from.scratch_off;
    rs.scratch_bytes = from.scratch_bytes;
    from.scratch_bytes = 0;
    rs.big_lane = true;
    rs.status = take_status(s);
    {
        std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
        after_small_steps(s);
        RT_HIP(hipEventRecord(rs.start, s->stream));
        resume_render_wave(*rs.resume, WaveYield{s->pinned_dev, s->yield_gen + 1}, s->stream);
        resolve_finishing(s, rs.finishing, s->stream, wave_counters(*rs.resume));
        RT_HIP(hipMemcpyAsync(rs.status, wave_counters(*rs.resume), kWaveCounterWords * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, s->stream));
        RT_HIP(hipEventRecord(rs.end, s->stream));
    }
    s->stats.steps++;
    s->stats.yields++;
    s->last_big_ms = s->now_ms();
    for (Step &o : s->inflight)
        if (!o.big_lane) o.overlapped = rs.overlapped = true;
    s->inflight.push_back(std::move(rs));
}

// A request whose result is where the caller wants it: its scene pin and arena range go back.  Called
// with s->m held.
void complete_request(rt_server *s, const std::shared_ptr<Request> &r, double done_ms) {
    if (r->state == State::Done) return;  // failed meanwhile (fail_all)
    {
        std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
        r->scene->users--;
    }
    s->arena.release(r->offset, r->bytes);
    r->state = State::Done;
    r->st.done_ms = done_ms;
    r->st.status = RT_OK;
    s->stats.completed++;
    auto it = std::find(s->active.begin(), s->active.end(), r);
    if (it != s->active.end()) s->active.erase(it);
    s->cv_done.notify_all();
    s->cv_work.notify_all();  // the freed range may admit the queue head; the worker may be waiting to stop
}

bool copies_out(const Request &r) {
    return r.req.out && !(r.req.kind == RT_REQ_TRACE && (r.req.trace_flags & RT_REQ_DEVICE_BUFFERS));
}

// Completes inflight[i] (its end event has fired).  Called with s->m held.  A lane's steps retire
// in order, so a request's last step retires after its others.
void retire_step(rt_server *s, size_t i, double done_ms) {
    Step step = std::move(s->inflight[i]);
    s->inflight.erase(s->inflight.begin() + (std::ptrdiff_t)i);
    if (step.resume && !wave_pass_complete(step.status)) resume_step(s, step);
    float ms = 0.0f, a = 0.0f, b = 0.0f;
    if (hipEventElapsedTime(&ms, step.start, step.end) == hipSuccess && ms > 0.0f &&
        hipEventElapsedTime(&a, s->t0_event, step.start) == hipSuccess &&
        hipEventElapsedTime(&b, s->t0_event, step.end) == hipSuccess) {
        s->stats.busy_ms += uncovered_ms(s->recent, a, b);
        s->recent.emplace_back(a, b);
        if (s->recent.size() > 64) s->recent.pop_front();
        // rates only from steps that had the device to themselves
        const double us = 1000.0 * ms;
        if (!step.overlapped && step.samples && !step.rays)
            s->sample_rate = 0.5 * s->sample_rate + 0.5 * (double)step.samples / us;
        if (!step.overlapped && step.rays && !step.samples)
            s->ray_rate = 0.5 * s->ray_rate + 0.5 * (double)step.rays / us;
    }
    s->stats.samples += step.samples;
    s->stats.rays += step.rays;
    for (auto &r : step.finishing) {
        if (copies_out(*r)) {  // completed by completer_main once the result is in the caller's buffer
            s->completing.push_back(r);
            s->cv_complete.notify_one();
        } else {
            complete_request(s, r, done_ms);
        }
    }
    if (step.scratch_bytes) s->scratch.release(step.scratch_off, step.scratch_bytes);
    s->event_pool.push_back(step.start);
    s->event_pool.push_back(step.end);
    s->cv_done.notify_all();
}

void fail_all(rt_server *s, int status, const std::string &msg);

// Copies finished requests' results to the callers' host buffers, oldest first, then completes them.
void completer_main(rt_server *s) {
    std::unique_lock<std::mutex> lk(s->m);
    hipError_t err = hipSetDevice(s->ctx->device);
    for (;;) {
        s->cv_complete.wait(lk, [&] { return !s->completing.empty() || s->completer_stop; });
        if (s->completing.empty()) break;  // asked to stop, nothing left
        std::shared_ptr<Request> r = s->completing.front();
        s->completing.pop_front();
        const bool render = r->req.kind == RT_REQ_RENDER;
        const uint64_t off = render ? r->offset : r->offset + Arena::round(32ull * r->req.nrays);
        const uint64_t n = render ? 16ull * r->req.width * r->req.height : 16ull * r->req.nrays;
        const void *src = arena_ptr(s, off);  // the range stays allocated until complete_request
        lk.unlock();
        if (err == hipSuccess && n) err = hipMemcpy(r->req.out, src, n, hipMemcpyDeviceToHost);
        lk.lock();
        if (err != hipSuccess) {
            fail_all(s, RT_ERR_HIP, std::string("copying a result to the host: ") + hipGetErrorString(err));
            s->completing.clear();
            s->cv_work.notify_all();
            continue;
        }
        complete_request(s, r, s->now_ms());
    }
}

void fail_all(rt_server *s, int status, const std::string &msg) {
    s->fatal = status;
    s->last_error = msg;
    auto fail = [&](std::shared_ptr<Request> &r) {
        if (r->state == State::Done) return;
        r->state = State::Done;
        r->st.status = status;
        r->st.done_ms = s->now_ms();
        s->stats.failed++;
    };
    for (auto &r : s->queue) fail(r);
    for (auto &r : s->active) fail(r);
    {
        std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
        for (auto &r : s->queue) r->scene->users--;
        for (auto &r : s->active) r->scene->users--;
    }
    s->queue.clear();
    s->active.clear();
    s->inflight.clear();
    s->cv_done.notify_all();
}

void worker_main(rt_server *s) {
    std::unique_lock<std::mutex> lk(s->m);
    try {
        RT_HIP(hipSetDevice(s->ctx->device));
        for (;;) {
            // retire the steps that have finished, in either lane
            for (size_t i = 0; i < s->inflight.size();) {
                if (hipEventQuery(s->inflight[i].end) == hipSuccess) retire_step(s, i, s->now_ms());
                else ++i;
            }
            admit(s);
            const bool serial = (s->cfg.flags & RT_SERVER_SERIAL) != 0;
            bool progress = false;
            if (serial) {
                progress = s->inflight.size() < s->cfg.pipeline && enqueue_step(s, true);
            } else {
                if (lane_inflight(s, false) < s->cfg.pipeline && enqueue_step(s, false)) {
                    progress = true;
                    if (yieldable_inflight(s)) s->yield_pending = true;
                }
                // A yieldable pass gives the device up to the small steps enqueued since it started
                // once it has run yield_min_ms (so a stream of small steps cannot starve it); it is
                // resumed after them (resume_step).
                if (s->yield_pending) {
                    if (!yieldable_inflight(s)) {
                        s->yield_pending = false;
                    } else if (yieldable_run_ms(s) >= s->yield_min_ms) {
                        request_yield(s);
                        s->yield_pending = false;
                    }
                }
                // While small requests keep arriving, the big lane runs one slice at a time, and none is
                // enqueued while a small step runs unless the big lane has waited a whole step budget
                // already.  Its slices are single yieldable passes; with RT_SERVER_NO_YIELD they are
                // short instead.
                const double now = s->now_ms();
                const bool small_recent = now - s->last_small_ms < 2.0 * s->cfg.step_us / 1000.0;
                const bool held = small_recent && lane_inflight(s, false) > 0 &&
                                  now - s->last_big_ms < s->cfg.step_us / 1000.0;
                const bool one = small_recent || yieldable_inflight(s);  // a yielded pass resumes before the next
                if (!held && lane_inflight(s, true) < (one ? 1u : s->cfg.pipeline) && enqueue_step(s, true)) {
                    s->last_big_ms = now;
                    progress = true;
                }
            }
            if (progress) continue;
            if (!s->inflight.empty()) {
                // Wait for the oldest small-lane step when one is in flight (it is short).  With only
                // device-filling steps in flight, or a yield to request, poll instead, so a request
                // submitted meanwhile starts at once on the small lane rather than after the
                // device-filling step.
                size_t w = 0;
                while (w + 1 < s->inflight.size() && s->inflight[w].big_lane && !serial) ++w;
                if (serial || (!s->inflight[w].big_lane && !s->yield_pending)) {
                    hipEvent_t e = s->inflight[w].end;
                    lk.unlock();
                    const hipError_t err = hipEventSynchronize(e);
                    lk.lock();
                    RT_HIP(err);
                } else {
                    s->cv_work.wait_for(lk, std::chrono::microseconds(50));
                }
                continue;
            }
            if (s->stop && s->queue.empty() && s->active.empty()) break;
            if (!s->queue.empty() && s->active.empty())  // cannot happen: an empty arena fits any admitted size
                throw Error(RT_ERR_BUDGET, "queued request does not fit the arena");
            s->cv_work.wait(lk);
        }
    } catch (const Error &e) {
        fail_all(s, e.status, e.what());
    } catch (const std::exception &e) {
        fail_all(s, RT_ERR_HIP, e.what());
    }
}

void destroy_streams(rt_server *s) {
    if (s->pinned) (void)hipHostFree(s->pinned);  // after the streams' work (the caller synchronised them)
    s->pinned = s->pinned_dev = nullptr;
    if (s->t0_event) (void)hipEventDestroy(s->t0_event);
    if (s->small_stream) (void)hipStreamDestroy(s->small_stream);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    s->t0_event = nullptr;
    s->small_stream = s->stream = nullptr;
}

}  // namespace
}  // namespace rtgpu

using namespace rtgpu;

extern "C" {

int rt_server_create(rt_context *ctx, const rt_server_config *cfg, rt_server **out) {
    if (!ctx || !out) return RT_ERR_INVALID;
    *out = nullptr;
    std::unique_ptr<rt_server> s;
    try {
        s = std::make_unique<rt_server>();
        s->ctx = ctx;
        if (cfg) s->cfg = *cfg;
        if (!s->cfg.arena_bytes) s->cfg.arena_bytes = 4ull << 30;
        if (!s->cfg.max_active) s->cfg.max_active = 256;
        if (!s->cfg.step_us) s->cfg.step_us = 50000;
        if (!s->cfg.pipeline) s->cfg.pipeline = 2;
        s->cfg.pipeline = std::min(s->cfg.pipeline, 8u);
        RT_REQUIRE((s->cfg.flags & ~(RT_SERVER_SERIAL | RT_SERVER_NO_YIELD)) == 0, "unknown server flags");
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        RT_HIP(hipSetDevice(ctx->device));
        s->arena_mem.alloc(s->cfg.arena_bytes);
        s->scratch_base = s->cfg.arena_bytes / 2 / Arena::kAlign * Arena::kAlign;
        s->arena.reset(s->scratch_base);
        s->scratch.reset(s->cfg.arena_bytes - s->scratch_base);
        s->segments.alloc(8);
        RT_HIP(hipMemset(s->segments.p, 0, 8));
        const size_t pinned_bytes = sizeof(uint32_t) * (4 + 4 * kStatusSlots);
        RT_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->pinned), pinned_bytes, hipHostMallocCoherent));
        std::memset(s->pinned, 0, pinned_bytes);
        RT_HIP(hipHostGetDevicePointer(reinterpret_cast<void **>(&s->pinned_dev), s->pinned, 0));
        // non-blocking: the server's work does not serialise with the legacy default stream
        RT_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        RT_HIP(hipStreamCreateWithFlags(&s->small_stream, hipStreamNonBlocking));
        if (const char *v = std::getenv("RTGPU_SERVER_SLICE_DIV")) s->slice_div = std::max(1.0, std::atof(v));
        if (const char *v = std::getenv("RTGPU_SERVER_SMALL_DIV")) s->small_div = std::max(1.0, std::atof(v));
        if (const char *v = std::getenv("RTGPU_SERVER_YIELD_MIN_MS")) s->yield_min_ms = std::max(0.0, std::atof(v));
        RT_HIP(hipEventCreate(&s->t0_event));
        RT_HIP(hipEventRecord(s->t0_event, s->stream));
        s->stats.arena_bytes = s->arena.size();
        s->t0 = Clock::now();
        s->worker = std::thread(worker_main, s.get());
        s->completer = std::thread(completer_main, s.get());
    } catch (const Error &e) {
        ctx->last_error = e.what();
        if (s) destroy_streams(s.get());
        return e.status;
    } catch (const std::exception &e) {
        ctx->last_error = e.what();
        if (s) destroy_streams(s.get());
        return RT_ERR_INVALID;
    }
    *out = s.release();
    return RT_OK;
}

int rt_server_destroy(rt_server *s) {
    if (!s) return RT_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();  // returns once every request has completed or failed
    {
        std::lock_guard<std::mutex> lk(s->m);
        s->completer_stop = true;
    }
    s->cv_complete.notify_all();
    if (s->completer.joinable()) s->completer.join();
    const int st = s->fatal;
    for (hipStream_t st : {s->stream, s->small_stream})
        if (st) (void)hipStreamSynchronize(st);
    for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
    destroy_streams(s);
    delete s;
    return st;
}

int rt_server_submit(rt_server *s, const rt_request *req, uint64_t *id) {
    if (!s || !req) return RT_ERR_INVALID;
    auto r = std::make_shared<Request>();
    r->req = *req;
    try {
        RT_REQUIRE(req->kind == RT_REQ_RENDER || req->kind == RT_REQ_TRACE, "unknown request kind");
        {
            std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
            r->scene = &s->ctx->scene(req->scene_id);
        }
        if (req->kind == RT_REQ_RENDER) {
            RT_REQUIRE((req->render_flags & ~(kRenderFlags | RT_REQ_OUT_ACCUM)) == 0, "unknown render flags");
            RT_REQUIRE(req->width > 0 && req->height > 0 && (uint64_t)req->width * req->height < (1ull << 31),
                       "bad image size");
            RT_REQUIRE(req->spp > 0, "a render request needs at least one sample per pixel");
            r->remaining = req->spp;
        } else {
            RT_REQUIRE((req->trace_flags & ~(kTraceFlags | RT_REQ_DEVICE_BUFFERS)) == 0, "unknown trace flags");
            RT_REQUIRE(req->nrays == 0 || req->rays, "rays is NULL");
            RT_REQUIRE(!(req->trace_flags & RT_REQ_DEVICE_BUFFERS) || req->nrays == 0 || req->out,
                       "a query on device buffers needs out");
            r->remaining = req->nrays;
        }
        {
            // a tree too deep for the configured stack is refused here rather than failing a step
            std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
            const bool render = req->kind == RT_REQ_RENDER;
            check_stack(*r->scene, resolve_cfg(s->ctx, *r->scene, render ? &s->cfg.render_cfg : &s->cfg.trace_cfg,
                                               !render && (req->trace_flags & RT_TRACE_ANY)));
        }
        r->bytes = request_bytes(*req, (s->cfg.flags & RT_SERVER_SERIAL) != 0);
    } catch (const Error &e) {
        std::lock_guard<std::mutex> lk(s->m);
        s->last_error = e.what();
        return e.status;
    }
    std::lock_guard<std::mutex> lk(s->m);
    if (s->fatal != RT_OK) return s->fatal;
    // the request's own buffers must fit its half of the arena, and one pass over its pixels
    // must fit the per-sample storage half, so every admitted request can make progress
    const uint64_t pass = req->kind == RT_REQ_RENDER && !(s->cfg.flags & RT_SERVER_SERIAL)
                              ? Arena::round(16ull * req->width * req->height) : 0;
    if (r->bytes > s->arena.size() || pass > s->scratch.size()) {
        s->last_error = "request needs " + std::to_string(r->bytes) + " + " + std::to_string(pass) +
                        " bytes; the arena's halves hold " + std::to_string(s->arena.size()) + " + " +
                        std::to_string(s->scratch.size());
        return RT_ERR_BUDGET;
    }
    if (r->remaining) {  // the scene stays resident until the request completes
        std::lock_guard<std::recursive_mutex> clk(s->ctx->mu);
        auto it = s->ctx->scenes.find(req->scene_id);  // looked up again: it may have been released meanwhile
        if (it == s->ctx->scenes.end()) {
            s->last_error = "unknown scene id " + std::to_string(req->scene_id);
            return RT_ERR_NOT_FOUND;
        }
        r->scene = it->second.get();
        r->scene->users++;
    }
    r->id = s->next_id++;
    r->st.id = r->id;
    r->st.device_bytes = r->bytes;
    r->st.submit_ms = s->now_ms();
    s->stats.submitted++;
    s->all[r->id] = r;
    if (r->remaining == 0) {  // an empty ray query completes at once
        r->state = State::Done;
        r->st.admit_ms = r->st.done_ms = r->st.submit_ms;
        r->st.status = RT_OK;
        s->stats.completed++;
        if (id) *id = r->id;
        s->cv_done.notify_all();
        return RT_OK;
    }
    s->queue.push_back(r);
    if (id) *id = r->id;
    s->cv_work.notify_all();
    return RT_OK;
}

int rt_server_wait(rt_server *s, uint64_t id, int timeout_ms, rt_request_stats *stats) {
    if (!s) return RT_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->m);
    auto it = s->all.find(id);
    if (it == s->all.end()) {
        s->last_error = "unknown request id " + std::to_string(id);
        return RT_ERR_NOT_FOUND;
    }
    std::shared_ptr<Request> r = it->second;
    auto done = [&] { return r->state == State::Done; };
    if (timeout_ms < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) return RT_ERR_BUSY;
    if (stats) *stats = r->st;
    return r->st.status;
}

int rt_server_drain(rt_server *s) {
    if (!s) return RT_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->m);
    s->cv_done.wait(lk, [&] { return (s->queue.empty() && s->active.empty()) || s->fatal != RT_OK; });
    return s->fatal;
}

int rt_server_stats_get(rt_server *s, rt_server_stats *out) {
    if (!s || !out) return RT_ERR_INVALID;
    unsigned long long segs = 0;
    if (hipMemcpy(&segs, s->segments.p, 8, hipMemcpyDeviceToHost) != hipSuccess) return RT_ERR_HIP;
    std::lock_guard<std::mutex> lk(s->m);
    *out = s->stats;
    out->segments = segs;
    out->arena_peak = s->arena.peak();
    out->wall_ms = s->now_ms();
    out->active = (uint32_t)s->active.size();
    out->queued = (uint32_t)s->queue.size();
    return RT_OK;
}

int rt_server_replay(rt_server *s, const rt_request *reqs, const double *arrival_ms, uint32_t n,
                     rt_request_stats *out) {
    if (!s || (n && (!reqs || !arrival_ms))) return RT_ERR_INVALID;
    std::vector<uint64_t> ids(n, 0);
    std::vector<int> status(n, RT_OK);
    double base;
    {
        std::lock_guard<std::mutex> lk(s->m);
        base = s->now_ms();
    }
    const Clock::time_point start = Clock::now();
    for (uint32_t i = 0; i < n; ++i) {
        std::this_thread::sleep_until(start + std::chrono::duration_cast<Clock::duration>(
                                                  std::chrono::duration<double, std::milli>(arrival_ms[i])));
        status[i] = rt_server_submit(s, &reqs[i], &ids[i]);
    }
    int result = RT_OK;
    for (uint32_t i = 0; i < n; ++i) {
        rt_request_stats st{};
        if (status[i] == RT_OK) status[i] = rt_server_wait(s, ids[i], -1, &st);
        if (status[i] != RT_OK && result == RT_OK) result = status[i];
        if (out) {
            st.id = ids[i];
            st.status = status[i];
            // open loop: latency runs from the scheduled arrival, not from when submit returned
            st.submit_ms = arrival_ms[i];
            st.admit_ms = st.admit_ms ? st.admit_ms - base : 0.0;
            st.done_ms = st.done_ms ? st.done_ms - base : 0.0;
            out[i] = st;
        }
    }
    return result;
}

const char *rt_server_last_error(const rt_server *s) { return s ? s->last_error.c_str() : "null server"; }

}  // extern "C"
