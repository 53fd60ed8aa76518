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
//
// One launch may render several images (jobs), each with its own scene, camera,
// image and sample count: the work items of all jobs are concatenated, each job
// starting on a chunk boundary, so a wave's chunk belongs to one job and is found
// with one wave-uniform search of the job table.  Since a pixel's samples depend only
// on its accumulator (the alpha carry), rendering spp = a and then spp = b into the
// same image equals rendering spp = a + b, which lets a scheduler hand out samples
// in slices.
//
// A small image cannot fill the device one pixel per lane, and a lane running a pixel's
// samples one after another sets the launch's length.  The carry entering sample s is a
// pure function of the image alpha, the pixel's seed and s (each sample's first draw is
// the next carry), so a split job gives every (pixel, sample) its own lane: the lane
// walks the carry chain to its sample (integer and conversion work only), traces that
// sample and stores its colour; a second pass adds the samples in order, which keeps the
// image bit-identical.  Split items trace their own first segment.
#include <algorithm>
#include <cmath>
#include <cstdlib>

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
    RenderJob job0;         // single-image launches
    const RenderJob *jobs;  // multi-image launches (device table)
    uint32_t njobs, nitems, refill, leaf_weight, shade_min;
    unsigned long long *segments;
    unsigned int *work;
};

enum : uint32_t { kIdle = 0, kTracing = 1, kResolved = 2 };

struct PixelState {
    uint32_t phase;
    uint32_t job;  // index into the job table (multi-image launches)
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
    const uint32_t *wide;  // the job's BVH (multi-image launches)
    const LeafRecord *rec;
};

template <bool MULTI> __device__ __forceinline__ const RenderJob &job_of(const RenderArgs &a, const PixelState &ps) {
    if constexpr (MULTI) return a.jobs[ps.job];
    else return a.job0;
}

template <bool MULTI> __device__ __forceinline__ const LeafRecord *rec_of(const RenderArgs &a, const PixelState &ps) {
    if constexpr (MULTI) return ps.rec;
    else return a.job0.rec;
}

template <bool MULTI> __device__ __forceinline__ const uint32_t *wide_of(const RenderArgs &a, const PixelState &ps) {
    if constexpr (MULTI) return ps.wide;
    else return a.job0.wide;
}

__device__ __forceinline__ void primary(const RenderJob &J, uint32_t x, uint32_t y, v3 &o, v3 &d) {
    v3 p00{J.cam[0], J.cam[1], J.cam[2]}, du{J.cam[3], J.cam[4], J.cam[5]};
    v3 dv{J.cam[6], J.cam[7], J.cam[8]}, org{J.cam[9], J.cam[10], J.cam[11]};
    v3 ps = (p00 + (float)x * du) + (float)y * dv;
    o = org;
    d = normalize(normalize(ps - org));  // getRay normalises; rayColor normalises again
}

template <bool MULTI>
__device__ __forceinline__ void begin_segment(const RenderArgs &a, PixelState &ps, Trav &t, uint32_t &segs) {
    if (ps.bounce == 0 && ps.pc_valid) {
        ps.hslot = ps.pc_slot;
        ps.ht = ps.pc_t;
        ps.phase = kResolved;
        return;
    }
    trav_init(t, ps.o, ps.d, 0.001f, 10000000.0f, job_of<MULTI>(a, ps).nwide);
    ps.phase = kTracing;
    segs++;
}

template <bool MULTI>
__device__ __forceinline__ void begin_sample(const RenderArgs &a, PixelState &ps, Trav &t, uint32_t &segs) {
    const RenderJob &J = job_of<MULTI>(a, ps);
    ps.st = ps.seed0 + f2u_sat(ps.acc[3] * 4294967294.0f);  // the literal is 2^32 in fp32
    ps.next = rng_next(ps.st);
    primary(J, ps.x, ps.y, ps.o, ps.d);
    ps.color = v3{0.0f, 0.0f, 0.0f};
    ps.atten = v3{1.0f, 1.0f, 1.0f};
    ps.bounce = 0;
    if (J.depth == 0) {
        ps.hslot = kEmpty;
        ps.phase = kResolved;
        return;
    }
    begin_segment<MULTI>(a, ps, t, segs);
}

// returns false when the pixel is finished
template <bool MULTI>
__device__ __forceinline__ bool end_sample(const RenderArgs &a, PixelState &ps, Trav &t, uint32_t &segs) {
    const RenderJob &J = job_of<MULTI>(a, ps);
    if (J.group) {  // split job: hand the sample to the in-order sum, then the item's next one
        J.samples[(size_t)ps.sample * (J.width * J.height) + ps.pix] =
            make_float4(ps.color.x, ps.color.y, ps.color.z, ps.next);
        ps.sample++;
        if (ps.sample < J.spp && ps.sample % J.group != 0) {
            ps.acc[3] = ps.next;
            begin_sample<MULTI>(a, ps, t, segs);
            return true;
        }
        ps.phase = kIdle;
        return false;
    }
    ps.acc[0] = ps.color.x + ps.acc[0];
    ps.acc[1] = ps.color.y + ps.acc[1];
    ps.acc[2] = ps.color.z + ps.acc[2];
    ps.acc[3] = ps.next;
    ps.sample++;
    if (ps.sample < J.spp) {
        begin_sample<MULTI>(a, ps, t, segs);
        return true;
    }
    reinterpret_cast<float4 *>(J.accum)[ps.pix] = make_float4(ps.acc[0], ps.acc[1], ps.acc[2], ps.acc[3]);
    ps.phase = kIdle;
    return false;
}

// emitted + scatter for a resolved segment (raytraceBVH.comp:94-115, 276-308)
template <bool MULTI>
__device__ __forceinline__ void shade(const RenderArgs &a, PixelState &ps, Trav &t, uint32_t &segs) {
    const RenderJob &J = job_of<MULTI>(a, ps);
    if (ps.bounce == 0 && !ps.pc_valid && !(J.flags & RT_RENDER_NO_PRIMARY_CACHE)) {
        ps.pc_slot = ps.hslot;
        ps.pc_t = ps.ht;
        ps.pc_valid = true;
    }
    if (ps.hslot == kEmpty) {
        const v3 background{0.0f, 0.0f, 0.0f};  // _BACKGROUND_COLOR
        ps.color = ps.color + mul(background, ps.atten);
        end_sample<MULTI>(a, ps, t, segs);
        return;
    }
    const LeafRecord r = rec_of<MULTI>(a, ps)[ps.hslot];
    const rt_material m = J.mats[J.slot_mat[ps.hslot]];
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
    const bool ref_mats = (J.flags & RT_RENDER_REFERENCE_MATERIALS) != 0;
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
        end_sample<MULTI>(a, ps, t, segs);
        return;
    }
    ps.atten = mul(ps.atten, att);
    ps.o = p;
    ps.d = nd;
    ps.bounce++;
    if (ps.bounce >= J.depth) {
        end_sample<MULTI>(a, ps, t, segs);
        return;
    }
    begin_segment<MULTI>(a, ps, t, segs);
}

// `local` is the item index within its job.  Pixel jobs: 64 items per 8x8 tile.  Split
// jobs: 64 items per group of `group` samples per tile, group-major within the tile, so
// neighbouring lanes trace the same samples of neighbouring pixels; `sample` is the
// item's first sample.
__device__ __forceinline__ bool item_to_pixel(const RenderJob &J, uint32_t local, uint32_t &x, uint32_t &y,
                                              uint32_t &sample) {
    uint32_t tile, within;
    if (J.group) {
        const uint32_t per_tile = 64u * ((J.spp + J.group - 1) / J.group);
        tile = local / per_tile;
        const uint32_t r = local - tile * per_tile;
        sample = (r >> 6) * J.group;
        within = r & 63u;
    } else {
        tile = local >> 6;
        within = local & 63u;
        sample = 0;
    }
    x = (tile % J.tiles_x) * 8u + (within & 7u);
    y = (tile / J.tiles_x) * 8u + (within >> 3);
    return x < J.width && y < J.height;
}

// Alpha carry entering sample `s` of a pixel whose image alpha is `a`: each sample seeds
// its RNG from the carry and its first draw becomes the next carry (raytraceBVH.comp:353-375).
__device__ __forceinline__ float carry_at(uint32_t seed0, float a, uint32_t s) {
    for (uint32_t k = 0; k < s; ++k) {
        uint32_t st = seed0 + f2u_sat(a * 4294967294.0f);
        a = rng_next(st);
    }
    return a;
}

// last job whose first item is <= g (wave-uniform arguments, so this runs on scalar loads)
__device__ __forceinline__ uint32_t find_job(const RenderJob *jobs, uint32_t njobs, uint32_t g) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].first_item <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int W, int S, bool MULTI>
__global__ __launch_bounds__(256) void k_render(RenderArgs a) {
    extern __shared__ uint32_t lds[];
    uint2 spill[kSpillCap];
    const StackMem m = stack_mem<S>(lds, spill);
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lt = lane ? (~0ull >> (64u - lane)) : 0ull;
    PixelState ps;
    ps.phase = kIdle;
    ps.job = 0;
    Trav t;
    t.have = false;
    t.top = 0;
    uint32_t segs = 0;
    bool exhausted = false;
    uint32_t q_next = 0, q_end = 0, q_job = 0;  // the wave's current chunk of work items (wave-uniform)
    for (;;) {
        const unsigned long long idle = __ballot(ps.phase == kIdle);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && nidle >= a.refill) {
            if (q_next == q_end) {
                const int leader = __ffsll((unsigned long long)idle) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) base = atomicAdd(a.work, kRenderChunk);
                base = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)base, leader));
                if (base >= a.nitems) {
                    exhausted = true;
                } else {
                    q_next = base;
                    q_end = (uint32_t)min((uint64_t)base + kRenderChunk, (uint64_t)a.nitems);
                    if constexpr (MULTI) q_job = find_job(a.jobs, a.njobs, base);
                }
            }
            const uint32_t take = exhausted ? 0u : min(nidle, q_end - q_next);
            const uint32_t r = (uint32_t)__popcll(idle & lt);
            const uint32_t item = q_next + r;
            q_next += take;
            if (ps.phase == kIdle && r < take) {
                ps.job = q_job;
                const RenderJob &J = job_of<MULTI>(a, ps);
                uint32_t x, y, sample;
                if (item_to_pixel(J, item - J.first_item, x, y, sample)) {
                    if constexpr (MULTI) {
                        ps.wide = J.wide;
                        ps.rec = J.rec;
                    }
                    ps.x = x;
                    ps.y = y;
                    ps.pix = y * J.width + x;
                    const float4 acc = reinterpret_cast<const float4 *>(J.accum)[ps.pix];
                    ps.acc[0] = acc.x; ps.acc[1] = acc.y; ps.acc[2] = acc.z; ps.acc[3] = acc.w;
                    ps.seed0 = rng_seed(x, y, J.random_state);
                    ps.sample = sample;
                    if (J.group) ps.acc[3] = carry_at(ps.seed0, acc.w, sample);  // this item's samples only
                    ps.pc_valid = false;
                    if (J.spp == 0) ps.phase = kIdle;
                    else begin_sample<MULTI>(a, ps, t, segs);
                }
            }
        }
        if (__ballot(ps.phase != kIdle) == 0ull) break;
        // Shading (RNG draws, sampling, scatter, the next segment's setup) is long and only a
        // few lanes finish a segment per iteration, so a wave shades only when at least
        // shade_min of its lanes hold a resolved segment, or when none is left tracing; the
        // resolved lanes wait meanwhile.  A lane's own sequence of segments is unchanged.
        const unsigned long long mres = __ballot(ps.phase == kResolved), mtr = __ballot(ps.phase == kTracing);
        if (mres != 0ull && (mtr == 0ull || (uint32_t)__popcll(mres) >= a.shade_min)) {
            // resolved segments may chain (cached primary hit, depth reached): bounded by spp*depth
            while (ps.phase == kResolved) shade<MULTI>(a, ps, t, segs);
        } else {
            const bool tracing = ps.phase == kTracing;
            const bool on_node = tracing && at_node(t), on_leaf = tracing && at_leaf(t);
            if (choose_leaf_phase(on_node, on_leaf, a.leaf_weight)) {
                if (on_leaf) leaf_step<S, false, false>(t, m, rec_of<MULTI>(a, ps));
            } else {
                if (on_node) node_step<W, S, false, false>(t, m, wide_of<MULTI>(a, ps));
            }
            if (tracing && trav_done(t)) {
                ps.hslot = t.bslot;
                ps.ht = t.bt;
                ps.phase = kResolved;
            }
        }
    }
    if (a.segments) {
        unsigned long long s = segs;
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == 0) atomicAdd(a.segments, s);
    }
}

// In-order sum of a split job's samples into its image: the per-pixel sequence of the
// reference's spp dispatches (rgb += sample colour, alpha = the sample's carry).  Samples
// are sample-major ([sample][pixel]) so both the writers and this pass are coalesced.
__device__ __forceinline__ void sum_job(const RenderJob &J) {
    if (!J.group) return;
    const uint32_t npix = J.width * J.height;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < npix; p += gridDim.x * blockDim.x) {
        float4 acc = reinterpret_cast<const float4 *>(J.accum)[p];
        for (uint32_t k = 0; k < J.spp; ++k) {
            const float4 v = J.samples[(size_t)k * npix + p];
            acc.x = v.x + acc.x;
            acc.y = v.y + acc.y;
            acc.z = v.z + acc.z;
            acc.w = v.w;
        }
        reinterpret_cast<float4 *>(J.accum)[p] = acc;
    }
}

template <bool MULTI> __global__ void k_sum_samples(RenderArgs a) {
    if constexpr (MULTI) sum_job(a.jobs[blockIdx.y]);
    else sum_job(a.job0);
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

template <int W, bool MULTI> void *pick_render_s(uint32_t s) {
    return s == 4 ? (void *)k_render<W, 4, MULTI> : s == 16 ? (void *)k_render<W, 16, MULTI> : (void *)k_render<W, 8, MULTI>;
}

void *pick_render(uint32_t width, uint32_t s, bool multi) {
    switch (width) {
        case 2: return multi ? pick_render_s<2, true>(s) : pick_render_s<2, false>(s);
        case 8: return multi ? pick_render_s<8, true>(s) : pick_render_s<8, false>(s);
        default: return multi ? pick_render_s<4, true>(s) : pick_render_s<4, false>(s);
    }
}

void launch_render_kernel(rt_context *ctx, RenderArgs &a, uint32_t bvh_width, bool multi, const LaunchCfg &c,
                          hipStream_t stream) {
    a.refill = c.refill;
    a.leaf_weight = c.leaf_weight;
    a.shade_min = c.shade_min;
    a.work = ctx->counter_slot();
    k_zero<<<1, 64, 0, stream>>>(a.work);
    const size_t lds = stack_lds_bytes(c.lds_stack, c.block);
    void *k = pick_render(bvh_width, c.lds_stack, multi);
    int per_cu = 0;
    RT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, (int)c.block, lds));
    const uint64_t waves_needed = (a.nitems + kRenderChunk - 1) / kRenderChunk;  // no more waves than chunks
    const uint64_t full = (uint64_t)std::max(1, per_cu) * (uint32_t)ctx->num_cus;
    const uint64_t wpb = c.block / 64u;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min(full, (waves_needed + wpb - 1) / wpb));
    void *args[] = {&a};
    RT_HIP(hipLaunchKernel(k, dim3(grid), dim3(c.block), args, lds, stream));
    RT_HIP(hipGetLastError());
}

// The in-order sum for the split jobs of a launch (max_pix: largest split image).
void launch_sum_samples(RenderArgs &a, bool multi, uint32_t max_pix, hipStream_t stream) {
    if (max_pix == 0) return;
    const uint32_t gx = std::min<uint32_t>((max_pix + 255) / 256, 4096);
    void *args[] = {&a};
    void *k = multi ? (void *)k_sum_samples<true> : (void *)k_sum_samples<false>;
    RT_HIP(hipLaunchKernel(k, dim3(gx, multi ? a.njobs : 1u), dim3(256), args, 0, stream));
    RT_HIP(hipGetLastError());
}

}  // namespace

void make_split(RenderJob &j, float4 *samples, uint32_t group) {
    RT_REQUIRE(samples, "split job without sample storage");
    RT_REQUIRE(group >= 1, "split group must hold at least one sample");
    const uint64_t groups = (j.spp + group - 1) / group;
    const uint64_t items = (uint64_t)j.tiles_x * ((j.height + 7) / 8) * 64u * groups;
    RT_REQUIRE(items < (1ull << 32) - kRenderChunk, "too many samples in one split launch");
    j.group = group;
    j.samples = samples;
    j.nitems = (uint32_t)items;  // each item traces its first segment once for its own samples
}

namespace {
uint64_t device_lanes(const rt_context *ctx) { return 1024ull * (uint64_t)ctx->num_cus; }  // ~resident lanes
}  // namespace

uint32_t split_group(const rt_context *ctx, uint32_t width, uint32_t height, uint32_t spp, uint32_t flags) {
    if (flags & RT_RENDER_NO_SPLIT || spp < 2) return 0;
    if (flags & RT_RENDER_SPLIT_SAMPLES) return 1;
    // Work items of `group` samples, enough of them (~32 per resident lane) that the launch's
    // drain -- lanes finishing their last, longest items -- is short; each item traces its
    // first segment once.  Whole pixels when that already holds, or when the per-sample
    // storage would be unreasonably large.
    uint64_t per_lane = 32;
    if (const char *e = std::getenv("RTGPU_ITEMS_PER_LANE")) per_lane = (uint64_t)std::max(1, std::atoi(e));
    const uint64_t samples = (uint64_t)width * height * spp;
    const uint64_t group = std::max<uint64_t>(1, samples / (per_lane * device_lanes(ctx)));
    if (group >= spp || 16ull * samples > (4ull << 30)) return 0;
    return (uint32_t)group;
}

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

RenderJob make_render_job(const Scene &sc, const rt_params &p, uint32_t width, uint32_t height, uint32_t spp,
                          uint32_t flags, float *accum) {
    RT_REQUIRE(accum, "accum is NULL");
    RT_REQUIRE(width > 0 && height > 0, "image size must be positive");
    RT_REQUIRE((uint64_t)width * height < (1ull << 31), "image too large");
    RenderJob j{};
    j.wide = sc.wide.as<uint32_t>();
    j.rec = sc.records.as<LeafRecord>();
    j.slot_mat = sc.slot_mat.as<uint32_t>();
    j.mats = sc.materials.as<rt_material>();
    j.accum = accum;
    j.nwide = sc.nwide;
    j.width = width;
    j.height = height;
    j.spp = spp;
    j.depth = p.maxRayTraceDepth;
    j.random_state = p.randomState;
    j.flags = flags;
    j.tiles_x = (width + 7) / 8;
    j.first_item = 0;
    j.nitems = j.tiles_x * ((height + 7) / 8) * 64u;
    j.samples = nullptr;
    j.group = 0;
    camera_setup(p, width, height, j.cam);
    return j;
}

void launch_render(rt_context *ctx, const Scene &sc, const rt_params &p, uint32_t width, uint32_t height, uint32_t spp,
                   uint32_t flags, float *accum, unsigned long long *segments, const LaunchCfg &c,
                   hipStream_t stream) {
    RenderJob j = make_render_job(sc, p, width, height, spp, flags, accum);
    if (spp == 0) return;
    check_stack(sc, c);
    float4 *samples = nullptr;
    const uint32_t group = split_group(ctx, width, height, spp, flags);
    const bool split = group != 0;
    if (split) {  // stream-ordered scratch from the device's memory pool, freed after the sum
        RT_HIP(hipMallocAsync(reinterpret_cast<void **>(&samples), 16ull * width * height * spp, stream));
        make_split(j, samples, group);
    }
    RenderArgs a{};
    a.job0 = j;
    a.jobs = nullptr;
    a.njobs = 1;
    a.nitems = j.nitems;
    a.segments = segments;
    launch_render_kernel(ctx, a, sc.width, false, c, stream);
    if (split) {
        launch_sum_samples(a, false, width * height, stream);
        RT_HIP(hipFreeAsync(samples, stream));
    }
}

uint64_t launch_render_jobs(rt_context *ctx, std::vector<RenderJob> &jobs, uint32_t bvh_width,
                            unsigned long long *segments, const LaunchCfg &c, hipStream_t stream) {
    uint64_t total = 0;
    uint32_t max_split_pix = 0;
    for (RenderJob &j : jobs) {
        total = (total + kRenderChunk - 1) / kRenderChunk * kRenderChunk;  // chunks never span jobs
        j.first_item = (uint32_t)total;
        total += j.spp ? j.nitems : 0u;  // a job with no samples gets no items
        if (j.group && j.spp) max_split_pix = std::max(max_split_pix, j.width * j.height);
    }
    if (total == 0) return 0;
    RT_REQUIRE(total < (1ull << 32) - kRenderChunk, "too many work items in one launch");
    int slot = 0;
    void *table = ctx->tables.stage(jobs.data(), sizeof(RenderJob) * jobs.size(), stream, slot);
    RenderArgs a{};
    a.jobs = static_cast<const RenderJob *>(table);
    a.njobs = (uint32_t)jobs.size();
    a.nitems = (uint32_t)total;
    a.segments = segments;
    launch_render_kernel(ctx, a, bvh_width, true, c, stream);
    launch_sum_samples(a, true, max_split_pix, stream);
    ctx->tables.commit(slot, stream);
    return total;
}

}  // namespace rtgpu
