// rt_internal.hpp -- host-side state behind the C ABI (context, resident scenes).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/rtgpu.h"
#include "rt_device.hpp"

namespace rtgpu {

// Error carried from internal code to the ABI boundary, where it becomes a status.
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

#define RT_HIP(call)                                                                                    \
    do {                                                                                                \
        hipError_t _e = (call);                                                                         \
        if (_e != hipSuccess)                                                                           \
            throw ::rtgpu::Error(_e == hipErrorOutOfMemory ? RT_ERR_NO_MEMORY : RT_ERR_HIP,             \
                                 std::string(#call) + ": " + hipGetErrorString(_e));                    \
    } while (0)

#define RT_REQUIRE(cond, msg)                                                                           \
    do {                                                                                                \
        if (!(cond)) throw ::rtgpu::Error(RT_ERR_INVALID, (msg));                                       \
    } while (0)

// Owning device allocation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        RT_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Launch configuration actually used by a kernel (after defaults are applied).
struct LaunchCfg {
    uint32_t block = 128;
    uint32_t lds_stack = 16;
    uint32_t persistent = 1;
    uint32_t refill = 16;
    uint32_t leaf_weight = 2;  // leaf phase when leaf lanes * weight >= node lanes
    uint32_t shade_min = 24;   // render: shade when this many lanes hold a resolved segment
};

// A scene resident in HBM.
struct Scene {
    uint32_t ntri = 0, nsph = 0, nmat = 0, nprim = 0;
    uint32_t width = 4, leaf_size = 1, flags = 0;
    uint32_t nwide = 0, wide_depth = 0, stack_bound = 0;
    float build_ms = 0.0f;
    DevBuf wide;       // nwide * 32*width bytes
    DevBuf records;    // nprim * LeafRecord
    DevBuf slot_prim;  // nprim * u32: leaf slot -> primitive id (RT_SPHERE_BIT for spheres)
    DevBuf slot_mat;   // nprim * u32: leaf slot -> material index
    DevBuf materials;  // nmat * rt_material
    DevBuf binary;     // (2N-1) * rt_bvh_node when RT_BUILD_KEEP_BINARY
    DevBuf tris;       // ntri * rt_triangle, world space (the refined ray-query results read them)
    DevBuf spheres;    // nsph * rt_sphere, world space
    bool has_tuned = false;
    LaunchCfg tuned_closest, tuned_any;
    uint32_t users = 0;  // server requests holding this scene
    // Streams whose work read this scene.  rt_scene_release waits for the context's use event of
    // each (recorded after every entry point's launches on that stream), not for the whole device:
    // a server's steps on other scenes keep running.  The event may stand for later work on the
    // same stream too, which only makes the wait longer.
    std::vector<hipStream_t> used_streams;
    bool unmarked_use = false;  // read by work that no use event covers: its release waits for the device

    Scene() = default;
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;

    // After an entry point's launches that read the scene on `stream`, with the result of
    // rt_context::record_use for that stream.
    void note_use(hipStream_t stream, int recorded);
    uint64_t device_bytes() const {
        return wide.bytes + records.bytes + slot_prim.bytes + slot_mat.bytes + materials.bytes + binary.bytes +
               tris.bytes + spheres.bytes;
    }
};

// One image of a (possibly multi-image) render launch: where its scene lives, its
// camera, and which slice of the launch's work items belongs to it.  Work items are
// pixels in 8x8 tiles; each job's items start at a multiple of kRenderChunk so a
// wave's chunk of items never spans two jobs.
struct RenderJob {
    const uint32_t *wide;
    const LeafRecord *rec;
    const uint32_t *slot_mat;
    const rt_material *mats;
    float *accum;
    float4 *samples;   // split jobs: per (pixel, sample) {colour, next carry}, [sample][pixel]
    uint32_t nwide, width, height, spp;
    uint32_t depth, random_state, flags, tiles_x;
    uint32_t first_item, nitems, group, _pad;  // group: samples per split item (0 = whole pixels)
    float cam[12];  // pixel00, deltaU, deltaV, origin
};
// Split jobs give every (pixel, group of `group` consecutive samples) its own work item; the
// samples are summed in order by a second pass.  `samples` must hold width*height*spp entries.
void make_split(RenderJob &j, float4 *samples, uint32_t group);
// Samples per split item for a whole-image render (0 = one lane per pixel).
uint32_t split_group(const rt_context *ctx, uint32_t width, uint32_t height, uint32_t spp, uint32_t flags);
constexpr uint32_t kRenderFlags = RT_RENDER_REFERENCE_MATERIALS | RT_RENDER_NO_PRIMARY_CACHE |
                                   RT_RENDER_SPLIT_SAMPLES | RT_RENDER_NO_SPLIT | RT_RENDER_WAVEFRONT |
                                   RT_RENDER_PERSISTENT_PATHS;
constexpr uint32_t kTraceFlags = RT_TRACE_ANY | RT_TRACE_FP32;
constexpr uint32_t kRenderChunk = 64;  // work items (one 8x8 tile) per atomic on the render work counter

// Small host->device tables (batch item lists, render job lists) that a launch reads.
// Each slot has a pinned host buffer and a device buffer; a slot is reused only after
// the launch that read it has finished, so launches need no host sync.  Completion is
// tracked on the ring's own stream: commit() marks the caller's stream, makes the ring
// stream wait for that mark and records the slot's `done` there.  The host only ever
// waits on events recorded on the ring stream, which lives as long as the ring, never on
// events of a caller's stream that may have been destroyed since.
class TableRing {
    static constexpr int kSlots = 128;
    struct Slot {
        void *host = nullptr;
        DevBuf dev;
        size_t cap = 0;
        hipEvent_t mark = nullptr;  // recorded on the caller's stream, waited on by the ring stream only
        hipEvent_t done = nullptr;  // recorded on the ring stream
        bool pending = false;
    };
    Slot slots_[kSlots];
    uint32_t next_ = 0;
    hipStream_t ring_ = nullptr;

public:
    TableRing() = default;
    TableRing(const TableRing &) = delete;
    TableRing &operator=(const TableRing &) = delete;
    ~TableRing() {
        if (ring_) (void)hipStreamSynchronize(ring_);
        for (Slot &s : slots_) {
            if (s.mark) (void)hipEventDestroy(s.mark);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.host) (void)hipHostFree(s.host);
        }
        if (ring_) (void)hipStreamDestroy(ring_);
    }
    // Enqueue a copy of `bytes` from `src` into a device table on `stream`.  Call
    // commit(slot, stream) after the launch that reads the table.
    void *stage(const void *src, size_t bytes, hipStream_t stream, int &slot) {
        if (!ring_) RT_HIP(hipStreamCreateWithFlags(&ring_, hipStreamNonBlocking));
        slot = (int)(next_++ % kSlots);
        Slot &s = slots_[slot];
        if (!s.done) {
            RT_HIP(hipEventCreateWithFlags(&s.mark, hipEventDisableTiming));
            RT_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
        }
        if (s.pending) RT_HIP(hipEventSynchronize(s.done));
        s.pending = false;
        if (s.cap < bytes) {
            size_t cap = 4096;
            while (cap < bytes) cap <<= 1;
            if (s.host) RT_HIP(hipHostFree(s.host));
            s.host = nullptr;
            s.cap = 0;
            RT_HIP(hipHostMalloc(&s.host, cap, hipHostMallocDefault));
            s.dev.alloc(cap);
            s.cap = cap;
        }
        std::memcpy(s.host, src, bytes);
        RT_HIP(hipMemcpyAsync(s.dev.p, s.host, bytes, hipMemcpyHostToDevice, stream));
        commit(slot, stream);  // covers the copy even if the launch is never enqueued
        return s.dev.p;
    }
    void commit(int slot, hipStream_t stream) {
        Slot &s = slots_[slot];
        RT_HIP(hipEventRecord(s.mark, stream));
        RT_HIP(hipStreamWaitEvent(ring_, s.mark, 0));
        RT_HIP(hipEventRecord(s.done, ring_));
        s.pending = true;
    }
};

}  // namespace rtgpu

namespace rtgpu {

// Device scratch of launches (split samples, path storage), kept for reuse.  A block is
// taken for the host-side enqueue of one launch sequence and given back with an event
// recorded on its stream after its last use: the same stream may take it again at once
// (stream order), another stream once the event has completed.  New blocks are allocated
// only when no free block is large enough; completed free blocks beyond kKeep bytes are
// freed.  The stream-ordered allocator's pools were avoided: with multi-GiB requests of
// varying size they stalled single launches for hundreds of milliseconds.
class ScratchCache {
  public:
    static constexpr uint64_t kKeep = 16ull << 30;
    static constexpr uint64_t kGrain = 64ull << 20;

    struct Block {
        DevBuf buf;
        hipEvent_t done = nullptr;
        hipStream_t last = nullptr;  // stream of the last use (valid when used)
        bool used = false, taken = false;
    };

    ~ScratchCache() {
        for (auto &b : blocks_)
            if (b->done) (void)hipEventDestroy(b->done);
    }

    Block *take(uint64_t bytes, hipStream_t st) {
        Block *best = nullptr;
        for (auto &b : blocks_) {
            if (b->taken || b->buf.bytes < bytes || !usable(*b, st)) continue;
            if (!best || b->buf.bytes < best->buf.bytes) best = b.get();
        }
        if (!best) {
            trim(bytes);
            auto b = std::make_unique<Block>();
            b->buf.alloc((bytes + kGrain - 1) / kGrain * kGrain);
            RT_HIP(hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
            best = b.get();
            blocks_.push_back(std::move(b));
        }
        best->taken = true;
        return best;
    }

    // Never throws (a lease gives its block back from a destructor).  A block whose event
    // cannot be recorded stays taken: its last use could never be known to have completed.
    void give(Block *b, hipStream_t st) noexcept {
        if (hipEventRecord(b->done, st) != hipSuccess) return;
        b->last = st;
        b->used = true;
        b->taken = false;
    }

  private:
    std::vector<std::unique_ptr<Block>> blocks_;

    static bool usable(const Block &b, hipStream_t st) {
        return !b.used || b.last == st || hipEventQuery(b.done) == hipSuccess;
    }
    // frees completed free blocks while the cache and the coming block exceed kKeep
    void trim(uint64_t coming) {
        uint64_t held = coming;
        for (auto &b : blocks_) held += b->buf.bytes;
        for (auto it = blocks_.begin(); it != blocks_.end() && held > kKeep;) {
            Block &b = **it;
            if (!b.taken && (!b.used || hipEventQuery(b.done) == hipSuccess)) {
                held -= b.buf.bytes;
                if (b.done) (void)hipEventDestroy(b.done);
                it = blocks_.erase(it);
            } else {
                ++it;
            }
        }
    }
};

// A cache block held for one scope of enqueues.
class ScratchLease {
  public:
    ScratchLease(ScratchCache &c, uint64_t bytes, hipStream_t st) : c_(c), st_(st), b_(c.take(bytes, st)) {}
    ~ScratchLease() { c_.give(b_, st_); }
    ScratchLease(const ScratchLease &) = delete;
    ScratchLease &operator=(const ScratchLease &) = delete;
    char *ptr() const { return b_->buf.as<char>(); }

  private:
    ScratchCache &c_;
    hipStream_t st_;
    ScratchCache::Block *b_;
};

}  // namespace rtgpu

struct rt_context {
    int device = 0;
    std::string last_error;
    uint32_t next_scene = 1;
    std::map<uint32_t, std::unique_ptr<rtgpu::Scene>> scenes;
    int num_cus = 256;
    // small device area for per-launch work counters (one slot per launch, round robin)
    rtgpu::DevBuf counters;
    uint32_t next_counter = 0;
    static constexpr uint32_t kCounterSlots = 4096;
    rtgpu::TableRing tables;
    // launch configuration chosen by rt_autotune_batch for fused multi-scene launches
    bool has_batch_tuned = false;
    rtgpu::LaunchCfg batch_closest, batch_any;
    // Held by every entry point for its duration and by a server's worker thread while
    // it enqueues a step, so a context may be shared by the caller and its servers.
    std::recursive_mutex mu;
    rtgpu::ScratchCache scratch;  // launch scratch (split samples, path storage)
    // the device address of this GPU's static scratch array (rt_render_noalloc), taken at rt_create
    char *noalloc_scratch = nullptr;

    // Per stream: an event recorded after the last entry point's launches on it (see record_use).
    std::vector<std::pair<hipStream_t, hipEvent_t>> use_events;
    static constexpr int kUseRecorded = 1, kUseCaptured = 0, kUseUnrecorded = -1;

    rt_context() = default;
    rt_context(const rt_context &) = delete;
    rt_context &operator=(const rt_context &) = delete;
    ~rt_context() {
        for (auto &u : use_events) (void)hipEventDestroy(u.second);
    }

    // Called once after an entry point has enqueued its launches on `stream` (the context lock
    // held).  kUseCaptured: the stream is being captured into a graph, whose caller orders the
    // graph's launches.  kUseUnrecorded: the stream has no event yet and the caller may not
    // create one (it must not allocate).
    int record_use(hipStream_t stream, bool may_create = true) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (stream && hipStreamIsCapturing(stream, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
            return kUseCaptured;
        hipEvent_t ev = nullptr;
        for (auto &u : use_events)
            if (u.first == stream) ev = u.second;
        if (!ev && !may_create) return kUseUnrecorded;
        if (!ev) {
            RT_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            use_events.emplace_back(stream, ev);
        }
        RT_HIP(hipEventRecord(ev, stream));
        return kUseRecorded;
    }
    // Blocks until the work that read `sc` has completed.
    void wait_unused(const rtgpu::Scene &sc) {
        if (sc.unmarked_use) RT_HIP(hipDeviceSynchronize());
        for (hipStream_t st : sc.used_streams)
            for (auto &u : use_events)
                if (u.first == st) RT_HIP(hipEventSynchronize(u.second));
    }

    rtgpu::Scene &scene(uint32_t id) {
        auto it = scenes.find(id);
        if (it == scenes.end()) throw rtgpu::Error(RT_ERR_NOT_FOUND, "unknown scene id " + std::to_string(id));
        return *it->second;
    }
    unsigned int *counter_slot() {
        uint32_t k = next_counter++ % kCounterSlots;
        return counters.as<unsigned int>() + 4u * k;
    }
};

namespace rtgpu {

inline void Scene::note_use(hipStream_t stream, int recorded) {
    if (recorded == rt_context::kUseCaptured) return;
    if (recorded == rt_context::kUseUnrecorded) {
        unmarked_use = true;
        return;
    }
    if (std::find(used_streams.begin(), used_streams.end(), stream) == used_streams.end())
        used_streams.push_back(stream);
}

// Runs f at the ABI boundary: holds the context lock, selects its device, and turns
// exceptions into a status code with the message kept for rt_last_error.
template <class F> inline int guarded(rt_context *ctx, F &&f) {
    std::unique_lock<std::recursive_mutex> lk;
    if (ctx) lk = std::unique_lock<std::recursive_mutex>(ctx->mu);
    try {
        if (ctx) {
            hipError_t e = hipSetDevice(ctx->device);
            if (e != hipSuccess) throw Error(RT_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
        }
        f();
        return RT_OK;
    } catch (const Error &e) {
        if (ctx) ctx->last_error = e.what();
        return e.status;
    } catch (const std::bad_alloc &) {
        if (ctx) ctx->last_error = "host allocation failed";
        return RT_ERR_NO_MEMORY;
    } catch (const std::exception &e) {
        if (ctx) ctx->last_error = e.what();
        return RT_ERR_INVALID;
    }
}

// build.hip
void build_scene(rt_context *ctx, Scene &sc, const rt_triangle *tris, uint32_t ntri, const rt_sphere *sph,
                 uint32_t nsph, const rt_material *mats, uint32_t nmat, const float *models, uint32_t nmodels,
                 const rt_bvh_node *nodes, uint32_t nnodes, const rt_build_options &opts);

// trace.hip
LaunchCfg resolve_cfg(const rt_context *ctx, const Scene &sc, const rt_launch_config *cfg, bool any);
// throws RT_ERR_OVERFLOW when the tree's worst-case stack depth exceeds the kernel's capacity
void check_stack(const Scene &sc, const LaunchCfg &cfg);
void launch_trace(rt_context *ctx, const Scene &sc, const rt_ray *rays, uint32_t n, rt_hit *hits,
                  uint32_t *counts, uint32_t flags, const LaunchCfg &cfg, hipStream_t stream);
// Closest-hit query whose hits carry leaf slots (for shading).  With n_dev set (a count produced
// on the device) it covers rays[0, min(*n_dev, n_max)) with the persistent kernel; without, it
// covers rays[0, n_max) with one thread per ray (camera rays).
void launch_trace_queue(rt_context *ctx, const Scene &sc, const rt_ray *rays, uint32_t n_max, const uint32_t *n_dev,
                        rt_hit *hits, const LaunchCfg &c, hipStream_t stream);
void launch_trace_batch(rt_context *ctx, const rt_batch_item *items, uint32_t nitems, uint32_t flags,
                        const LaunchCfg &cfg, hipStream_t stream);
struct TraceItem {
    const Scene *sc;
    const rt_ray *rays;
    rt_hit *hits;
    uint32_t n;
};
// One persistent launch per BVH width present over all items' rays; no host sync.
void launch_trace_items(rt_context *ctx, const std::vector<TraceItem> &items, uint32_t flags, const LaunchCfg &cfg,
                        hipStream_t stream);

// render.hip
void launch_accum_clear(float *accum, uint32_t npix, hipStream_t stream);
// Bounce-by-bounce rendering of one job into its image (see RT_RENDER_WAVEFRONT), with the
// path storage from `scratch` (at least wave_scratch_bytes(width * height)) or, when it is
// NULL, from the context's scratch cache.
uint64_t wave_scratch_bytes(uint64_t paths);
// Pixel-samples from which an image fills the device on its own and is rendered bounce by
// bounce (by rt_render, and by the server, which then runs it alone)
constexpr uint64_t kWaveMinSamples = 1ull << 22;
// rt_render's choice of schedule: bounce by bounce (true) or one persistent launch
bool wave_schedule(uint32_t width, uint32_t height, uint32_t spp, uint32_t flags);
uint64_t per_pass_limit(uint64_t npix);  // samples of an npix-pixel image per pass
// With `primary` (npix hits, caller-owned), the pixels' primary hits are kept there: traced by this call
// unless `primary_ready`, then read by every pass.  A render cut into several calls traces them once
// (camera rays carry no jitter, so every sample of a pixel starts with the same hit).
//
// Yielding (`yield` set, `resume` receives the pass): a render of one pass whose last bounces run
// in the finish kernel can give the device up early.  Once *yield->flag reaches yield->yield_at,
// the finish kernel's waves take no more queue entries and exit when their paths end; the pass's
// sum (and a resolve gated on wave_counters) then does nothing, and resume_render_wave runs the
// entries left, as often as it yields in turn.  The image is the same however often it yields.
// When the render is not one such pass, *resume is left empty and nothing yields.
struct WaveResume;
struct WaveYield {
    uint32_t *flag;     // yield requests, in fine-grained host memory the device reads
    uint32_t yield_at;  // the launch yields once *flag >= yield_at
};
void launch_render_wave(rt_context *ctx, const Scene &sc, const RenderJob &j, unsigned long long *segments,
                        const LaunchCfg &c, hipStream_t stream, char *scratch, uint64_t scratch_bytes,
                        rt_hit *primary = nullptr, bool primary_ready = false, const WaveYield *yield = nullptr,
                        std::shared_ptr<WaveResume> *resume = nullptr);
// Runs the entries a yielded finish launch left (its counter continues where that launch stopped),
// then the pass's gated sum.  The pass's storage must be as the first launch left it.
void resume_render_wave(const WaveResume &r, const WaveYield &yield, hipStream_t stream);
// The pass's device counters (kWaveCounterWords words): its two queues' lengths, the finish
// kernel's work counter and that counter's value at a yield.  wave_pass_complete tells from a copy
// of them whether the pass is complete (else it yielded and must be resumed).
constexpr uint32_t kWaveCounterWords = 4;
const uint32_t *wave_counters(const WaveResume &r);
bool wave_pass_complete(const uint32_t *counters);
// With `scratch` set, every buffer beyond the inputs and outputs comes from it (scratch_bytes long):
// bounce-by-bounce renders run in as many passes as it needs (RT_ERR_BUDGET when one sample of the
// image does not fit), and persistent launches whose per-sample storage does not fit run whole pixels
// per lane (the same image).  Without it, storage comes from the context's scratch cache.
void launch_render(rt_context *ctx, const Scene &sc, const rt_params &p, uint32_t width, uint32_t height,
                   uint32_t spp, uint32_t flags, float *accum, unsigned long long *segments,
                   const LaunchCfg &cfg, hipStream_t stream, char *scratch = nullptr, uint64_t scratch_bytes = 0);
// noalloc.hip: the device address of the static scratch array on the current device (hipGetSymbolAddress)
void *noalloc_scratch_address();
uint64_t noalloc_scratch_bytes();
RenderJob make_render_job(const Scene &sc, const rt_params &p, uint32_t width, uint32_t height, uint32_t spp,
                          uint32_t flags, float *accum);
// One launch over all jobs (all scenes must have BVH width `bvh_width`); returns the work items launched.
uint64_t launch_render_jobs(rt_context *ctx, std::vector<RenderJob> &jobs, uint32_t bvh_width,
                            unsigned long long *segments, const LaunchCfg &cfg, hipStream_t stream);
// With `gate` (a yieldable pass's counters, WaveResume::counters), nothing is written unless that pass is complete.
void launch_resolve(const float *accum, uint32_t npix, uint32_t spp, float *out, hipStream_t stream,
                    const uint32_t *gate = nullptr);

// host camera (raytraceBVH.comp:49-81); out: pixel00, deltaU, deltaV, origin
void camera_setup(const rt_params &p, uint32_t width, uint32_t height, float out[12]);

}  // namespace rtgpu
