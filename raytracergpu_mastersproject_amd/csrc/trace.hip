// trace.hip -- ray-query kernels (closest-hit / any-hit) over the resident wide BVH.
//
// Two schedules:
//   persistent: a grid sized to fill every CU once; each wave takes chunks of rays
//     from a global counter, one atomic per chunk (ChunkSchedule: smaller chunks for
//     the last rays).  When at least `refill` lanes of a wave are idle they take the
//     next rays of the wave's chunk (ballot + popcount + lane prefix), so finished
//     lanes are refilled while long rays keep traversing.
//   direct: one lane per ray (a launch-tuner candidate; the camera rays of renders).
// The batched variant runs many scenes' ray batches in one persistent launch.
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "rt_internal.hpp"
#include "traverse.hpp"

namespace rtgpu {
namespace {

struct SceneView {
    const uint32_t *wide;
    const LeafRecord *rec;
    const uint32_t *slot_prim;
    const rt_triangle *tris;  // world-space primitives, for the refined results
    const rt_sphere *sph;
    uint32_t nwide;
    uint32_t _pad;
};

struct TraceArgs {
    SceneView sc;  // sc.slot_prim == nullptr: hits carry the leaf slot instead of the primitive id
    const rt_ray *rays;
    rt_hit *hits;
    uint32_t *counts;
    uint32_t n;
    uint32_t refill;
    unsigned int *work;
    uint32_t leaf_weight;
    const uint32_t *n_dev;  // when set, the ray count is min(*n_dev, n)
};

// Rays are read once and hits written once per launch: non-temporal accesses keep them from
// displacing the BVH nodes and leaf records in L2 and the MALL (refined closest-hit query on
// 4M triangles: 5.10 against 5.13 ms).
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef unsigned int nt_u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void store_hit_nt(rt_hit *hits, uint32_t i, uint4 h) {
    const nt_u4 v = {h.x, h.y, h.z, h.w};
    __builtin_nontemporal_store(v, reinterpret_cast<nt_u4 *>(hits) + i);
}

__device__ __forceinline__ void load_ray(const rt_ray *rays, uint32_t i, v3 &o, float &tmin, v3 &d, float &tmax) {
    const nt_f4 *p = reinterpret_cast<const nt_f4 *>(rays + i);
    const nt_f4 a = __builtin_nontemporal_load(p), b = __builtin_nontemporal_load(p + 1);
    o = v3{a.x, a.y, a.z};
    tmin = a.w;
    d = v3{b.x, b.y, b.z};
    tmax = b.w;
}

__device__ __forceinline__ void store_hit(const SceneView &sc, const Trav &t, rt_hit *hits, uint32_t *counts,
                                          uint32_t i) {
    uint4 h;
    if (t.bslot == kEmpty) h = make_uint4(RT_NO_HIT, __float_as_uint(t.bt), 0u, 0u);
    else h = make_uint4(sc.slot_prim ? sc.slot_prim[t.bslot] : t.bslot, __float_as_uint(t.bt), __float_as_uint(t.bu),
                        __float_as_uint(t.bv));
    store_hit_nt(hits, i, h);
    if (counts) reinterpret_cast<uint2 *>(counts)[i] = make_uint2(t.nbox, t.ntri);
}

template <int W, int S, bool ANY, bool COUNT>
__global__ __launch_bounds__(256) void k_trace_direct(TraceArgs a) {
    extern __shared__ uint32_t lds[];
    uint2 spill[kSpillCap];
    const StackMem m = stack_mem<S>(lds, spill);
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (a.n_dev ? min(*a.n_dev, a.n) : a.n)) return;
    v3 o, d;
    float tmin, tmax;
    load_ray(a.rays, i, o, tmin, d, tmax);
    Trav t;
    trav_init(t, o, d, tmin, tmax, a.sc.nwide);
    trav_run<W, S, ANY, COUNT>(t, m, a.sc.wide, a.sc.rec, a.leaf_weight);
    store_hit(a.sc, t, a.hits, a.counts, i);
}

// The stored result of a finished ray, with t, u, v of a hit refined in double as k_refine_hits
// refines them (the same operations on the same fp32 inputs, so the same bits).
__device__ __forceinline__ void store_refined(const SceneView &sc, const Trav &t, rt_hit *hits, uint32_t *counts,
                                              uint32_t i) {
    uint4 h;
    if (t.bslot == kEmpty) {
        h = make_uint4(RT_NO_HIT, __float_as_uint(t.bt), 0u, 0u);
    } else {
        const uint32_t prim = sc.slot_prim[t.bslot];
        float tt = t.bt, u = t.bu, v = t.bv;
        if (prim & RT_SPHERE_BIT) {
            const float4 c = *reinterpret_cast<const float4 *>(sc.sph + (prim & ~RT_SPHERE_BIT));
            const float r = sc.sph[prim & ~RT_SPHERE_BIT].radius;
            refine_sphere(dvec(c.x, c.y, c.z), (double)r, dvec(t.o.x, t.o.y, t.o.z), dvec(t.d.x, t.d.y, t.d.z), tt);
        } else {
            const float4 *tp = reinterpret_cast<const float4 *>(sc.tris + prim);
            const float4 p0 = tp[0], p1 = tp[1], p2 = tp[2];
            refine_triangle(dvec(p0.x, p0.y, p0.z), dvec(p1.x, p1.y, p1.z), dvec(p2.x, p2.y, p2.z),
                            dvec(t.o.x, t.o.y, t.o.z), dvec(t.d.x, t.d.y, t.d.z), tt, u, v);
        }
        h = make_uint4(prim, __float_as_uint(tt), __float_as_uint(u), __float_as_uint(v));
    }
    store_hit_nt(hits, i, h);
    if (counts) reinterpret_cast<uint2 *>(counts)[i] = make_uint2(t.nbox, t.ntri);
}

// REFINE: store each hit with t, u, v refined in double (store_refined) instead of the fp32 values.
template <int W, int S, bool ANY, bool COUNT, bool REFINE>
__global__ __launch_bounds__(256) void k_trace_persistent(TraceArgs a) {
    extern __shared__ uint32_t lds[];
    uint2 spill[kSpillCap];
    const StackMem m = stack_mem<S>(lds, spill);
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lt = lane ? (~0ull >> (64u - lane)) : 0ull;
    int64_t rid = -1;
    // REFINE: a finished ray's result is stored at the wave's next refill (or at the end), where the
    // lanes that finished since the last refill refine together; refining as each lane finishes would
    // run the double-precision code on almost every iteration.  Without refinement a lane stores at once.
    int64_t done = -1;
    bool exhausted = false;
    uint32_t q_next = 0, q_end = 0;  // the wave's current chunk of ray indices (wave-uniform)
    const uint32_t n = a.n_dev ? min(*a.n_dev, a.n) : a.n;
    const ChunkSchedule<uint32_t> sched(n, gridDim.x * (blockDim.x / 64u), true);
    Trav t;
    t.have = false;
    t.top = 0;
    const auto store = [&](uint32_t i) {
        if constexpr (REFINE) store_refined(a.sc, t, a.hits, a.counts, i);
        else store_hit(a.sc, t, a.hits, a.counts, i);
    };
    for (;;) {
        const unsigned long long idle = __ballot(rid < 0);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && nidle >= a.refill) {
            if (done >= 0) {
                store((uint32_t)done);
                done = -1;
            }
            if (q_next == q_end) {  // one atomic per chunk keeps the shared counter off the critical path
                const int leader = __ffsll((unsigned long long)idle) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) base = atomicAdd(a.work, 1u);
                base = (uint32_t)__shfl((int)base, leader);
                uint64_t lo, hi;
                if (!sched.range(base, lo, hi)) exhausted = true;
                else { q_next = (uint32_t)lo; q_end = (uint32_t)hi; }
            }
            if (!exhausted) {
                const uint32_t take = min(nidle, q_end - q_next);
                const uint32_t r = (uint32_t)__popcll(idle & lt);
                if (rid < 0 && r < take) {
                    rid = (int64_t)(q_next + r);
                    v3 o, d;
                    float tmin, tmax;
                    load_ray(a.rays, (uint32_t)rid, o, tmin, d, tmax);
                    trav_init(t, o, d, tmin, tmax, a.sc.nwide);
                }
                q_next += take;
            }
        }
        if (__ballot(rid >= 0) == 0ull) break;  // every lane idle and the queue is drained
        const bool on_node = rid >= 0 && at_node(t), on_leaf = rid >= 0 && at_leaf(t);
        if (choose_leaf_phase(on_node, on_leaf, a.leaf_weight)) {
            if (on_leaf) leaf_step<S, ANY, COUNT>(t, m, a.sc.rec);
        } else {
            if (on_node) node_step<W, S, ANY, COUNT>(t, m, a.sc.wide);
        }
        if (rid >= 0 && trav_done(t)) {
            if constexpr (REFINE) done = rid;
            else store((uint32_t)rid);
            rid = -1;
        }
    }
    if (done >= 0) store((uint32_t)done);
}

// ---------------------------------------------------------------------------
// Refined results: t, u, v of every hit recomputed in double precision from the hit
// primitive's stored fp32 geometry and the fp32 ray (refine_triangle / refine_sphere).
// The primitive id, decided by the fp32 traversal, is untouched; so is a result whose
// recomputation has no solution in double (a zero determinant, a negative discriminant).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void refine_one(const rt_triangle *tris, const rt_sphere *sph, const rt_ray *rays,
                                           rt_hit *hits, uint32_t i) {
    uint4 *hp = reinterpret_cast<uint4 *>(hits) + i;
    uint4 h = *hp;
    if (h.x == RT_NO_HIT) return;
    v3 o, d;
    float tmin, tmax;
    load_ray(rays, i, o, tmin, d, tmax);
    float t = __uint_as_float(h.y), u = __uint_as_float(h.z), v = __uint_as_float(h.w);
    bool ok;
    if (h.x & RT_SPHERE_BIT) {
        const float4 c = *reinterpret_cast<const float4 *>(sph + (h.x & ~RT_SPHERE_BIT));
        const float r = sph[h.x & ~RT_SPHERE_BIT].radius;
        ok = refine_sphere(dvec(c.x, c.y, c.z), (double)r, dvec(o.x, o.y, o.z), dvec(d.x, d.y, d.z), t);
    } else {
        const float4 *tp = reinterpret_cast<const float4 *>(tris + h.x);
        const float4 a = tp[0], b = tp[1], c = tp[2];
        ok = refine_triangle(dvec(a.x, a.y, a.z), dvec(b.x, b.y, b.z), dvec(c.x, c.y, c.z), dvec(o.x, o.y, o.z),
                             dvec(d.x, d.y, d.z), t, u, v);
    }
    if (ok) *hp = make_uint4(h.x, __float_as_uint(t), __float_as_uint(u), __float_as_uint(v));
}

__global__ __launch_bounds__(256) void k_refine_hits(const rt_ray *rays, rt_hit *hits, uint32_t n,
                                                     const rt_triangle *tris, const rt_sphere *sph) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) refine_one(tris, sph, rays, hits, i);
}

// ---------------------------------------------------------------------------
// batched: many (scene, rays) items in one persistent launch.  A global ray
// index runs over the concatenation of all items; item boundaries are found by
// binary search in an exclusive prefix of item sizes held in the argument block.
// ---------------------------------------------------------------------------
struct BatchItemDev {
    SceneView sc;
    const rt_ray *rays;
    rt_hit *hits;
    uint64_t first;  // global index of the item's first ray
};

struct BatchArgs {
    const BatchItemDev *items;
    uint32_t nitems;
    uint32_t refill;
    uint64_t total;
    unsigned long long *work;
    uint32_t leaf_weight;
};

// the last item in [lo, hi) whose first ray is at or before g
__device__ __forceinline__ uint32_t find_item(const BatchItemDev *items, uint32_t lo, uint32_t hi, uint64_t g) {
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (items[mid].first <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// REFINE: store each hit with t, u, v refined in double (store_refined).
template <int W, int S, bool ANY, bool REFINE>
__global__ __launch_bounds__(256) void k_trace_batch(BatchArgs a) {
    extern __shared__ uint32_t lds[];
    uint2 spill[kSpillCap];
    const StackMem m = stack_mem<S>(lds, spill);
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lt = lane ? (~0ull >> (64u - lane)) : 0ull;
    bool active = false, exhausted = false;
    bool pending = false;  // REFINE: a finished ray (item, local) whose result waits for the wave's next refill
    uint32_t item = 0, local = 0;
    uint64_t q_next = 0, q_end = 0;  // the wave's current chunk of global ray indices
    uint32_t c_lo = 0, c_hi = 1;     // the items the chunk's rays belong to: [c_lo, c_hi)
    // the lane's scene, kept in registers: a table load per traversal step would add its
    // latency to every step
    const uint32_t *wide = nullptr;
    const LeafRecord *rec = nullptr;
    const ChunkSchedule<uint64_t> sched(a.total, gridDim.x * (blockDim.x / 64u), false);
    Trav t;
    t.have = false;
    t.top = 0;
    const auto store = [&]() {
        const BatchItemDev &it = a.items[item];
        if constexpr (REFINE) store_refined(it.sc, t, it.hits, nullptr, local);
        else store_hit(it.sc, t, it.hits, nullptr, local);
    };
    for (;;) {
        const unsigned long long idle = __ballot(!active);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && nidle >= a.refill) {
            if (pending) {  // as in k_trace_persistent: the lanes that finished refine together
                store();
                pending = false;
            }
            if (q_next == q_end) {
                const int leader = __ffsll((unsigned long long)idle) - 1;
                unsigned long long base = 0;
                if ((int)lane == leader) base = atomicAdd(a.work, 1ull);
                base = __shfl(base, leader);
                if (!sched.range(base, q_next, q_end)) {
                    exhausted = true;
                } else {  // chunks are short: usually one item, found once per chunk
                    c_lo = find_item(a.items, 0u, a.nitems, q_next);
                    c_hi = find_item(a.items, c_lo, a.nitems, q_end - 1u) + 1u;
                }
            }
            const uint32_t take = exhausted ? 0u : (uint32_t)min((uint64_t)nidle, q_end - q_next);
            const uint32_t r = (uint32_t)__popcll(idle & lt);
            const uint64_t g = q_next + r;
            q_next += take;
            if (!active && r < take) {
                item = find_item(a.items, c_lo, c_hi, g);
                local = (uint32_t)(g - a.items[item].first);
                wide = a.items[item].sc.wide;
                rec = a.items[item].sc.rec;
                v3 o, d;
                float tmin, tmax;
                load_ray(a.items[item].rays, local, o, tmin, d, tmax);
                trav_init(t, o, d, tmin, tmax, a.items[item].sc.nwide);
                active = true;
            }
        }
        if (__ballot(active) == 0ull) break;
        const bool on_node = active && at_node(t), on_leaf = active && at_leaf(t);
        if (choose_leaf_phase(on_node, on_leaf, a.leaf_weight)) {
            if (on_leaf) leaf_step<S, ANY, false>(t, m, rec);
        } else {
            if (on_node) node_step<W, S, ANY, false>(t, m, wide);
        }
        if (active && trav_done(t)) {
            if constexpr (REFINE) pending = true;
            else store();
            active = false;
        }
    }
    if (pending) store();
}

__global__ void k_zero_counter(unsigned int *c) {
    if (threadIdx.x == 0) { c[0] = 0u; c[1] = 0u; }
}

SceneView view_of(const Scene &sc) {
    return SceneView{sc.wide.as<uint32_t>(), sc.records.as<LeafRecord>(), sc.slot_prim.as<uint32_t>(),
                     sc.tris.as<rt_triangle>(), sc.spheres.as<rt_sphere>(), sc.nwide, 0u};
}

template <int W, int S> struct Kernels {
    template <bool ANY> static void *direct(bool count) {
        return count ? (void *)k_trace_direct<W, S, ANY, true> : (void *)k_trace_direct<W, S, ANY, false>;
    }
    template <bool ANY, bool REFINE> static void *persistent(bool count) {
        return count ? (void *)k_trace_persistent<W, S, ANY, true, REFINE>
                     : (void *)k_trace_persistent<W, S, ANY, false, REFINE>;
    }
    static void *get(int which, bool any, bool count) {
        if (which == 0) return any ? direct<true>(count) : direct<false>(count);
        if (which == 1) return any ? persistent<true, false>(count) : persistent<false, false>(count);
        if (which == 3) return persistent<false, true>(count);  // closest hit only
        if (which == 2) return any ? (void *)k_trace_batch<W, S, true, false> : (void *)k_trace_batch<W, S, false, false>;
        return any ? (void *)k_trace_batch<W, S, true, true> : (void *)k_trace_batch<W, S, false, true>;
    }
};

template <int W> void *pick(uint32_t s, int which, bool any, bool count) {
    switch (s) {
        case 4: return Kernels<W, 4>::get(which, any, count);
        case 16: return Kernels<W, 16>::get(which, any, count);
        default: return Kernels<W, 8>::get(which, any, count);
    }
}

// which: 0 one lane per ray, 1 persistent, 2 persistent over many scenes (never counts), 3 persistent
// storing refined hits, 4 persistent over many scenes storing refined hits
void *pick_kernel(uint32_t width, uint32_t s, int which, bool any, bool count) {
    switch (width) {
        case 2: return pick<2>(s, which, any, count);
        case 8: return pick<8>(s, which, any, count);
        default: return pick<4>(s, which, any, count);
    }
}

uint32_t persistent_grid(rt_context *ctx, const void *kernel, uint32_t block, size_t lds) {
    int per_cu = 0;
    RT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block, lds));
    return (uint32_t)std::max(1, per_cu) * (uint32_t)ctx->num_cus;
}

}  // namespace

void check_stack(const Scene &sc, const LaunchCfg &c) {
    if (sc.stack_bound > c.lds_stack + kSpillCap)
        throw Error(RT_ERR_OVERFLOW, "traversal stack bound " + std::to_string(sc.stack_bound) +
                                         " exceeds the kernel's capacity " + std::to_string(c.lds_stack + kSpillCap));
}

LaunchCfg resolve_cfg(const rt_context *ctx, const Scene &sc, const rt_launch_config *cfg, bool any) {
    (void)ctx;
    LaunchCfg c = sc.has_tuned ? (any ? sc.tuned_any : sc.tuned_closest) : LaunchCfg{};
    if (cfg) {
        if (cfg->block) c.block = cfg->block;
        if (cfg->lds_stack) c.lds_stack = cfg->lds_stack;
        c.persistent = cfg->persistent;
        if (cfg->refill) c.refill = cfg->refill;
    }
    RT_REQUIRE(c.block == 64 || c.block == 128 || c.block == 256, "block must be 64, 128 or 256");
    RT_REQUIRE(c.lds_stack == 4 || c.lds_stack == 8 || c.lds_stack == 16, "lds_stack must be 4, 8 or 16");
    c.refill = std::min<uint32_t>(std::max<uint32_t>(c.refill, 1u), 64u);
    if (const char *lw = std::getenv("RTGPU_LEAF_WEIGHT")) c.leaf_weight = (uint32_t)std::max(1, std::atoi(lw));
    if (const char *sm = std::getenv("RTGPU_SHADE_MIN")) c.shade_min = (uint32_t)std::min(64, std::max(1, std::atoi(sm)));
    return c;
}

void launch_trace(rt_context *ctx, const Scene &sc, const rt_ray *rays, uint32_t n, rt_hit *hits, uint32_t *counts,
                  uint32_t flags, const LaunchCfg &c, hipStream_t stream) {
    if (n == 0) return;
    RT_REQUIRE(rays && hits, "rays/hits must be device pointers");
    check_stack(sc, c);
    const bool any = (flags & RT_TRACE_ANY) != 0, refine = (flags & RT_TRACE_FP32) == 0;
    // Closest-hit persistent launches refine in the kernel, at each wave's refills: measured on one box,
    // 5.11-5.14 ms per 16M-ray query against 5.19 ms with the separate pass.  Any-hit rays end sooner
    // and feel the refining kernel's registers (90 VGPRs, 5 waves per SIMD instead of 6) more: 3,270
    // against 3,410 Mrays/s, so they keep the separate pass.
    const bool refine_in_kernel = refine && !any && c.persistent;
    TraceArgs a{view_of(sc), rays, hits, counts, n, c.refill, nullptr, c.leaf_weight, nullptr};
    const size_t lds = stack_lds_bytes(c.lds_stack, c.block);
    if (!c.persistent) {
        void *k = pick_kernel(sc.width, c.lds_stack, 0, any, counts != nullptr);
        const uint32_t grid = (uint32_t)(((uint64_t)n + c.block - 1) / c.block);
        void *args[] = {&a};
        RT_HIP(hipLaunchKernel(k, dim3(grid), dim3(c.block), args, lds, stream));
    } else {
        a.work = ctx->counter_slot();
        k_zero_counter<<<1, 64, 0, stream>>>(a.work);
        void *k = pick_kernel(sc.width, c.lds_stack, refine_in_kernel ? 3 : 1, any, counts != nullptr);
        // no more lanes than rays: a small batch would otherwise launch the whole device's worth of
        // waves, nearly all of which find no chunk and exit
        const uint32_t grid = (uint32_t)std::min<uint64_t>(persistent_grid(ctx, k, c.block, lds),
                                                           ((uint64_t)n + c.block - 1) / c.block);
        void *args[] = {&a};
        RT_HIP(hipLaunchKernel(k, dim3(grid), dim3(c.block), args, lds, stream));
    }
    if (refine && !refine_in_kernel)
        k_refine_hits<<<(uint32_t)(((uint64_t)n + 255u) / 256u), 256, 0, stream>>>(rays, hits, n, a.sc.tris, a.sc.sph);
    RT_HIP(hipGetLastError());
}

void launch_trace_queue(rt_context *ctx, const Scene &sc, const rt_ray *rays, uint32_t n_max, const uint32_t *n_dev,
                        rt_hit *hits, const LaunchCfg &c, hipStream_t stream) {
    if (n_max == 0) return;
    check_stack(sc, c);
    SceneView v = view_of(sc);
    v.slot_prim = nullptr;  // hits carry leaf slots
    TraceArgs a{v, rays, hits, nullptr, n_max, c.refill, ctx->counter_slot(), c.leaf_weight, n_dev};
    const size_t lds = stack_lds_bytes(c.lds_stack, c.block);
    if (!n_dev) {
        // A count known on the host means camera rays: neighbouring lanes trace neighbouring
        // pixels, and about two thirds of them leave the scene at once.  One thread per ray lets
        // such waves retire and new ones start (0.45 vs 0.69 ms for the 1080p frame's primaries);
        // refilling lanes in a persistent wave mixes rays at different depths instead.
        void *k = pick_kernel(sc.width, c.lds_stack, 0, false, false);
        void *args[] = {&a};
        RT_HIP(hipLaunchKernel(k, dim3((uint32_t)(((uint64_t)n_max + c.block - 1) / c.block)), dim3(c.block), args, lds, stream));
        RT_HIP(hipGetLastError());
        return;
    }
    k_zero_counter<<<1, 64, 0, stream>>>(a.work);
    void *k = pick_kernel(sc.width, c.lds_stack, 1, false, false);
    const uint32_t grid = persistent_grid(ctx, k, c.block, lds);
    void *args[] = {&a};
    RT_HIP(hipLaunchKernel(k, dim3(grid), dim3(c.block), args, lds, stream));
    RT_HIP(hipGetLastError());
}

void launch_trace_items(rt_context *ctx, const std::vector<TraceItem> &items, uint32_t flags, const LaunchCfg &c,
                        hipStream_t stream) {
    // one launch per BVH width present, each over that width's items in order
    for (uint32_t width : {2u, 4u, 8u}) {
        std::vector<BatchItemDev> hv;
        uint64_t total = 0;
        for (const TraceItem &it : items) {
            if (it.sc->width != width) continue;
            RT_REQUIRE(it.n == 0 || (it.rays && it.hits), "batch item with NULL rays/hits");
            check_stack(*it.sc, c);
            hv.push_back(BatchItemDev{view_of(*it.sc), it.rays, it.hits, total});
            total += it.n;
        }
        if (total == 0) continue;
        int slot = 0;
        void *table = ctx->tables.stage(hv.data(), sizeof(BatchItemDev) * hv.size(), stream, slot);
        unsigned long long *work = reinterpret_cast<unsigned long long *>(ctx->counter_slot());
        k_zero_counter<<<1, 64, 0, stream>>>(reinterpret_cast<unsigned int *>(work));
        BatchArgs a{static_cast<const BatchItemDev *>(table), (uint32_t)hv.size(), c.refill, total, work,
                    c.leaf_weight};
        const bool any = (flags & RT_TRACE_ANY) != 0;
        const size_t lds = stack_lds_bytes(c.lds_stack, c.block);
        void *k = pick_kernel(width, c.lds_stack, (flags & RT_TRACE_FP32) ? 2 : 4, any, false);
        const uint64_t chunks = (total + kChunk - 1) / kChunk, wpb = c.block / 64u;
        const uint32_t grid = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>(persistent_grid(ctx, k, c.block, lds), (chunks + wpb - 1) / wpb));
        void *args[] = {&a};
        RT_HIP(hipLaunchKernel(k, dim3(grid), dim3(c.block), args, lds, stream));
        RT_HIP(hipGetLastError());
        ctx->tables.commit(slot, stream);
    }
}

void launch_trace_batch(rt_context *ctx, const rt_batch_item *items, uint32_t nitems, uint32_t flags,
                        const LaunchCfg &c, hipStream_t stream) {
    if (nitems == 0) return;
    RT_REQUIRE(items, "items is NULL");
    std::vector<TraceItem> v(nitems);
    for (uint32_t i = 0; i < nitems; ++i)
        v[i] = TraceItem{&ctx->scene(items[i].scene_id), items[i].rays, items[i].hits, items[i].nrays};
    launch_trace_items(ctx, v, flags, c, stream);
}

}  // namespace rtgpu
