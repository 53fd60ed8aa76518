// render.hip -- path tracing of raytraceBVH.comp (compute stage S2) on the wide BVH.
//
// The reference dispatches one sample per pixel per dispatch, spp times, with
// an image barrier between dispatches (RaytracerBVH.cpp:1025-1050); the
// samples of one pixel form a chain through the image's alpha channel (the RNG
// carry).  Pixels are independent, so here one lane runs all spp samples of a
// pixel in order, keeping the accumulator in registers: the image is read and
// written once per pixel instead of once per sample, with bit-identical
// results.  Lanes are persistent and take new pixels (8x8 tiles, so a wave's
// primary rays are coherent) as soon as they finish; each iteration of the
// outer loop advances every active lane by one traversal step.
//
// Primary rays have no jitter (raytraceBVH.comp:338, defocus angle 0), so the
// first segment of every sample of a pixel is the same ray with the same
// closest hit; it is traced once per pixel and reused (RT_RENDER_NO_PRIMARY_CACHE
// disables this).
#include <algorithm>
#include <cmath>

#include "rt_internal.hpp"
#include "traverse.hpp"

namespace rtgpu {

void camera_setup(const rt_params &p, uint32_t width, uint32_t height, float out[12]) {
    const float Wf = (float)width, Hf = (float)height;
    const float focal = 10.0f;  // _FOCAL_DISTANCE
    float aspect = Wf / Hf;
    float theta = p.verticalFOV * 0.017453292519943295f;  // radians()
    float h = (float)std::tan((double)(theta / 2.0f));
    float vh = 2.0f * h * focal;
    float vw = vh * aspect;
    v3 pos{p.camPos[0], p.camPos[1], p.camPos[2]};
    v3 at{p.camLookAt[0], p.camLookAt[1], p.camLookAt[2]};
    v3 up{p.camUpDir[0], p.camUpDir[1], p.camUpDir[2]};
    v3 w = normalize(pos - at);
    v3 u = normalize(cross(up, w));
    v3 v = cross(w, u);
    v3 vpu = vw * u;
    v3 vpv = vh * (-v);
    v3 du = vpu / Wf, dv = vpv / Hf;
    v3 ul = ((pos - focal * w) - vpu / 2.0f) - vpv / 2.0f;
    v3 p00 = ul + 0.5f * (du + dv);
    const v3 outs[4] = {p00, du, dv, pos};
    for (int k = 0; k < 4; ++k) {
        out[3 * k + 0] = outs[k].x;
        out[3 * k + 1] = outs[k].y;
        out[3 * k + 2] = outs[k].z;
    }
}

namespace {

struct RenderArgs {
    const uint32_t *wide;
    const LeafRecord *rec;
    const uint32_t *slot_mat;
    const rt_material *mats;
    uint32_t nwide;
    uint32_t width, height, spp, depth, random_state, flags, refill, leaf_weight;
    uint32_t tiles_x, nitems;
    float cam[12];
    float *accum;
    unsigned long long *segments;
    unsigned int *work;
};

enum : uint32_t { kIdle = 0, kTracing = 1, kResolved = 2 };
constexpr uint32_t kRenderChunk = 128;  // pixels (two 8x8 tiles) per atomic on the work counter

struct PixelState {
    uint32_t phase;
    uint32_t pix;  // y*width + x
    uint32_t x, y;
    uint32_t seed0, st;
    uint32_t sample, bounce;
    float acc[4];
    float next;
    v3 color, atten, o, d;
    uint32_t hslot;  // result of the current segment
    float ht;
    uint32_t pc_slot;  // cached first segment
    float pc_t;
    bool pc_valid;
};

__device__ __forceinline__ void primary(const RenderArgs &a, uint32_t x, uint32_t y, v3 &o, v3 &d) {
    v3 p00{a.cam[0], a.cam[1], a.cam[2]}, du{a.cam[3], a.cam[4], a.cam[5]};
    v3 dv{a.cam[6], a.cam[7], a.cam[8]}, org{a.cam[9], a.cam[10], a.cam[11]};
    v3 ps = (p00 + (float)x * du) + (float)y * dv;
    o = org;
    d = normalize(normalize(ps - org));  // getRay normalises; rayColor normalises again
}

__device__ __forceinline__ void begin_segment(const RenderArgs &a, PixelState &ps, Trav &t, uint32_t &segs) {
    if (ps.bounce == 0 && ps.pc_valid) {
        ps.hslot = ps.pc_slot;
        ps.ht = ps.pc_t;
        ps.phase = kResolved;
        return;
    }
    trav_init(t, ps.o, ps.d, 0.001f, 10000000.0f, a.nwide);
    ps.phase = kTracing;
    segs++;
}

__device__ __forceinline__ void begin_sample(const RenderArgs &a, PixelState &ps, Trav &t, uint32_t &segs) {
    ps.st = ps.seed0 + f2u_sat(ps.acc[3] * 4294967294.0f);  // the literal is 2^32 in fp32
    ps.next = rng_next(ps.st);
    primary(a, ps.x, ps.y, ps.o, ps.d);
    ps.color = v3{0.0f, 0.0f, 0.0f};
    ps.atten = v3{1.0f, 1.0f, 1.0f};
    ps.bounce = 0;
    if (a.depth == 0) {
        ps.hslot = kEmpty;
        ps.phase = kResolved;
        return;
    }
    begin_segment(a, ps, t, segs);
}

// returns false when the pixel is finished
__device__ __forceinline__ bool end_sample(const RenderArgs &a, PixelState &ps, Trav &t, uint32_t &segs) {
    ps.acc[0] = ps.color.x + ps.acc[0];
    ps.acc[1] = ps.color.y + ps.acc[1];
    ps.acc[2] = ps.color.z + ps.acc[2];
    ps.acc[3] = ps.next;
    ps.sample++;
    if (ps.sample < a.spp) {
        begin_sample(a, ps, t, segs);
        return true;
    }
    reinterpret_cast<float4 *>(a.accum)[ps.pix] = make_float4(ps.acc[0], ps.acc[1], ps.acc[2], ps.acc[3]);
    ps.phase = kIdle;
    return false;
}

// emitted + scatter for a resolved segment (raytraceBVH.comp:94-115, 276-308)
__device__ __forceinline__ void shade(const RenderArgs &a, PixelState &ps, Trav &t, uint32_t &segs) {
    if (ps.bounce == 0 && !ps.pc_valid && !(a.flags & RT_RENDER_NO_PRIMARY_CACHE)) {
        ps.pc_slot = ps.hslot;
        ps.pc_t = ps.ht;
        ps.pc_valid = true;
    }
    if (ps.hslot == kEmpty) {
        const v3 background{0.0f, 0.0f, 0.0f};  // _BACKGROUND_COLOR
        ps.color = ps.color + mul(background, ps.atten);
        end_sample(a, ps, t, segs);
        return;
    }
    const LeafRecord r = a.rec[ps.hslot];
    const rt_material m = a.mats[a.slot_mat[ps.hslot]];
    const v3 alb{m.albedo[0], m.albedo[1], m.albedo[2]};
    const v3 p = ps.o + ps.ht * ps.d;
    v3 n;
    if (is_sphere_record(r)) n = (p - v3{r.r0.x, r.r0.y, r.r0.z}) / r.r0.w;
    else n = v3{r.r1.w, r.r2.w, r.r3.w};
    const int bf = dot(ps.d, n) > 0.0f ? 1 : 0;
    const v3 normal = (float)(1 - 2 * bf) * n;
    const v3 emitted = m.materialType == RT_LIGHT ? alb : v3{0.0f, 0.0f, 0.0f};
    ps.color = ps.color + mul(emitted, ps.atten);
    bool scattered = false;
    v3 att{0.0f, 0.0f, 0.0f}, nd{0.0f, 0.0f, 0.0f};
    const bool ref_mats = (a.flags & RT_RENDER_REFERENCE_MATERIALS) != 0;
    if (m.materialType == RT_DIFFUSE) {
        att = alb;
        v3 ruv = normalize(random_in_unit_sphere(ps.st));
        nd = normalize(normal + ruv);
        scattered = true;
    } else if (m.materialType == RT_METALLIC && !ref_mats) {
        att = alb;
        v3 refl = reflect(ps.d, normal);
        v3 rs = random_in_unit_sphere(ps.st);
        nd = normalize(refl + m.albedo[3] * rs);
        scattered = dot(nd, normal) > 0.0f;
    } else if (m.materialType == RT_DIELECTRIC && !ref_mats) {
        att = v3{1.0f, 1.0f, 1.0f};
        const float ior = m.albedo[3];
        const float ri = bf ? ior : (1.0f / ior);
        const float ct = gmin(dot(-ps.d, normal), 1.0f);
        const float stt = sqrtf(1.0f - ct * ct);
        const bool cannot = ri * stt > 1.0f;
        if (cannot || schlick(ct, ri) > rng_next(ps.st)) nd = reflect(ps.d, normal);
        else nd = refract(ps.d, normal, ri);
        nd = normalize(nd);
        scattered = true;
    }
    if (!scattered) {
        end_sample(a, ps, t, segs);
        return;
    }
    ps.atten = mul(ps.atten, att);
    ps.o = p;
    ps.d = nd;
    ps.bounce++;
    if (ps.bounce >= a.depth) {
        end_sample(a, ps, t, segs);
        return;
    }
    begin_segment(a, ps, t, segs);
}

__device__ __forceinline__ bool item_to_pixel(const RenderArgs &a, uint32_t item, uint32_t &x, uint32_t &y) {
    const uint32_t tile = item >> 6, within = item & 63u;
    x = (tile % a.tiles_x) * 8u + (within & 7u);
    y = (tile / a.tiles_x) * 8u + (within >> 3);
    return x < a.width && y < a.height;
}

template <int W, int S>
__global__ __launch_bounds__(256) void k_render(RenderArgs a) {
    extern __shared__ uint2 lds[];
    uint2 spill[kSpillCap];
    StackMem m{lds + threadIdx.x, blockDim.x, spill};
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lt = lane ? (~0ull >> (64u - lane)) : 0ull;
    PixelState ps;
    ps.phase = kIdle;
    Trav t;
    t.have = false;
    t.top = 0;
    uint32_t segs = 0;
    bool exhausted = false;
    uint32_t q_next = 0, q_end = 0;  // the wave's current chunk of work items (wave-uniform)
    for (;;) {
        const unsigned long long idle = __ballot(ps.phase == kIdle);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && nidle >= a.refill) {
            if (q_next == q_end) {
                const int leader = __ffsll((unsigned long long)idle) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) base = atomicAdd(a.work, kRenderChunk);
                base = (uint32_t)__shfl((int)base, leader);
                if (base >= a.nitems) exhausted = true;
                else { q_next = base; q_end = (uint32_t)min((uint64_t)base + kRenderChunk, (uint64_t)a.nitems); }
            }
            const uint32_t take = exhausted ? 0u : min(nidle, q_end - q_next);
            const uint32_t r = (uint32_t)__popcll(idle & lt);
            const uint32_t item = q_next + r;
            q_next += take;
            if (ps.phase == kIdle && r < take) {
                uint32_t x, y;
                if (item_to_pixel(a, item, x, y)) {
                    ps.x = x;
                    ps.y = y;
                    ps.pix = y * a.width + x;
                    const float4 acc = reinterpret_cast<const float4 *>(a.accum)[ps.pix];
                    ps.acc[0] = acc.x; ps.acc[1] = acc.y; ps.acc[2] = acc.z; ps.acc[3] = acc.w;
                    ps.seed0 = rng_seed(x, y, a.random_state);
                    ps.sample = 0;
                    ps.pc_valid = false;
                    if (a.spp == 0) ps.phase = kIdle;
                    else begin_sample(a, ps, t, segs);
                }
            }
        }
        if (__ballot(ps.phase != kIdle) == 0ull) break;
        const bool tracing = ps.phase == kTracing;
        const bool on_node = tracing && at_node(t), on_leaf = tracing && at_leaf(t);
        if (choose_leaf_phase(on_node, on_leaf, a.leaf_weight)) {
            if (on_leaf) leaf_step<S, false>(t, m, a.rec);
        } else {
            if (on_node) node_step<W, S, false>(t, m, a.wide);
        }
        if (tracing && trav_done(t)) {
            ps.hslot = t.bslot;
            ps.ht = t.bt;
            ps.phase = kResolved;
        }
        // resolved segments may chain (cached primary hit, depth reached): bounded by spp*depth
        while (ps.phase == kResolved) shade(a, ps, t, segs);
    }
    if (a.segments) {
        unsigned long long s = segs;
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) atomicAdd(a.segments, s);
    }
}

__global__ void k_accum_clear(float4 *accum, uint32_t npix) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x)
        accum[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
}

// SingleTriangleFullScreen.frag:13-21
__global__ void k_resolve(const float4 *accum, uint32_t npix, uint32_t spp, float4 *out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x) {
        float4 a = accum[i];
        float s = (float)spp;
        float r = gmin(gmax(sqrtf(a.x / s), 0.0f), 1.0f);
        float g = gmin(gmax(sqrtf(a.y / s), 0.0f), 1.0f);
        float b = gmin(gmax(sqrtf(a.z / s), 0.0f), 1.0f);
        out[i] = make_float4(r, g, b, 1.0f);
    }
}

__global__ void k_zero(unsigned int *c) {
    if (threadIdx.x == 0) c[0] = 0u;
}

void *pick_render(uint32_t width, uint32_t s) {
    switch (width) {
        case 2: return s == 4 ? (void *)k_render<2, 4> : s == 16 ? (void *)k_render<2, 16> : (void *)k_render<2, 8>;
        case 8: return s == 4 ? (void *)k_render<8, 4> : s == 16 ? (void *)k_render<8, 16> : (void *)k_render<8, 8>;
        default: return s == 4 ? (void *)k_render<4, 4> : s == 16 ? (void *)k_render<4, 16> : (void *)k_render<4, 8>;
    }
}

}  // namespace

void launch_accum_clear(float *accum, uint32_t npix, hipStream_t stream) {
    if (npix == 0) return;
    RT_REQUIRE(accum, "accum is NULL");
    uint32_t grid = std::min<uint32_t>((npix + 255) / 256, 65536);
    k_accum_clear<<<grid, 256, 0, stream>>>(reinterpret_cast<float4 *>(accum), npix);
    RT_HIP(hipGetLastError());
}

void launch_resolve(const float *accum, uint32_t npix, uint32_t spp, float *out, hipStream_t stream) {
    if (npix == 0) return;
    RT_REQUIRE(accum && out, "accum/out is NULL");
    uint32_t grid = std::min<uint32_t>((npix + 255) / 256, 65536);
    k_resolve<<<grid, 256, 0, stream>>>(reinterpret_cast<const float4 *>(accum), npix, spp,
                                        reinterpret_cast<float4 *>(out));
    RT_HIP(hipGetLastError());
}

void launch_render(rt_context *ctx, const Scene &sc, const rt_params &p, uint32_t width, uint32_t height, uint32_t spp,
                   uint32_t flags, float *accum, unsigned long long *segments, const LaunchCfg &c,
                   hipStream_t stream) {
    RT_REQUIRE(accum, "accum is NULL");
    RT_REQUIRE(width > 0 && height > 0, "image size must be positive");
    RT_REQUIRE((uint64_t)width * height < (1ull << 31), "image too large");
    if (spp == 0) return;
    if (sc.stack_bound > c.lds_stack + kSpillCap)
        throw Error(RT_ERR_OVERFLOW, "traversal stack bound exceeds the kernel's capacity");
    RenderArgs a{};
    a.wide = sc.wide.as<uint32_t>();
    a.rec = sc.records.as<LeafRecord>();
    a.slot_mat = sc.slot_mat.as<uint32_t>();
    a.mats = sc.materials.as<rt_material>();
    a.nwide = sc.nwide;
    a.width = width;
    a.height = height;
    a.spp = spp;
    a.depth = p.maxRayTraceDepth;
    a.random_state = p.randomState;
    a.flags = flags;
    a.refill = c.refill;
    a.leaf_weight = c.leaf_weight;
    a.tiles_x = (width + 7) / 8;
    const uint64_t items = (uint64_t)a.tiles_x * ((height + 7) / 8) * 64u;
    a.nitems = (uint32_t)items;
    camera_setup(p, width, height, a.cam);
    a.accum = accum;
    a.segments = segments;
    a.work = ctx->counter_slot();
    k_zero<<<1, 64, 0, stream>>>(a.work);
    const size_t lds = sizeof(uint2) * c.lds_stack * c.block;
    void *k = pick_render(sc.width, c.lds_stack);
    int per_cu = 0;
    RT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, (int)c.block, lds));
    uint32_t grid = (uint32_t)std::max(1, per_cu) * (uint32_t)ctx->num_cus;
    void *args[] = {&a};
    RT_HIP(hipLaunchKernel(k, dim3(grid), dim3(c.block), args, lds, stream));
    RT_HIP(hipGetLastError());
}

}  // namespace rtgpu
