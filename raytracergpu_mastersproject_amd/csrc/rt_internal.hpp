// rt_internal.hpp -- host-side state behind the C ABI (context, resident scenes).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
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
    uint32_t lds_stack = 8;
    uint32_t persistent = 1;
    uint32_t refill = 16;
    uint32_t leaf_weight = 2;  // leaf phase when leaf lanes * weight >= node lanes
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
    bool has_tuned = false;
    LaunchCfg tuned_closest, tuned_any;
    uint64_t device_bytes() const {
        return wide.bytes + records.bytes + slot_prim.bytes + slot_mat.bytes + materials.bytes + binary.bytes;
    }
};

struct rt_context_impl;

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
    std::mutex mu;  // guards scenes/counters for the server's worker thread

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

// build.hip
void build_scene(rt_context *ctx, Scene &sc, const rt_triangle *tris, uint32_t ntri, const rt_sphere *sph,
                 uint32_t nsph, const rt_material *mats, uint32_t nmat, const float *models, uint32_t nmodels,
                 const rt_bvh_node *nodes, uint32_t nnodes, const rt_build_options &opts);

// trace.hip
LaunchCfg resolve_cfg(const rt_context *ctx, const Scene &sc, const rt_launch_config *cfg, bool any);
void launch_trace(rt_context *ctx, const Scene &sc, const rt_ray *rays, uint32_t n, rt_hit *hits,
                  uint32_t *counts, uint32_t flags, const LaunchCfg &cfg, hipStream_t stream);
void launch_trace_batch(rt_context *ctx, const rt_batch_item *items, uint32_t nitems, uint32_t flags,
                        const LaunchCfg &cfg, hipStream_t stream);

// render.hip
void launch_accum_clear(float *accum, uint32_t npix, hipStream_t stream);
void launch_render(rt_context *ctx, const Scene &sc, const rt_params &p, uint32_t width, uint32_t height,
                   uint32_t spp, uint32_t flags, float *accum, unsigned long long *segments,
                   const LaunchCfg &cfg, hipStream_t stream);
void launch_resolve(const float *accum, uint32_t npix, uint32_t spp, float *out, hipStream_t stream);

// host camera (raytraceBVH.comp:49-81); out: pixel00, deltaU, deltaV, origin
void camera_setup(const rt_params &p, uint32_t width, uint32_t height, float out[12]);

}  // namespace rtgpu
