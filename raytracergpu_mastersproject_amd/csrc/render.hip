// render.hip -- path tracing of raytraceBVH.comp (compute stage S2) on the wide BVH.
//
// The reference dispatches one sample per pixel per dispatch, spp times, with an image
// barrier between dispatches (RaytracerBVH.cpp:1025-1050); the samples of one pixel form a
// chain through the image's alpha channel (the RNG carry).  Two schedules reproduce it bit
// for bit; every path performs the same operations in the same order under both.
//
// Bounce by bounce (images of 2^22+ pixel-samples, and RT_RENDER_WAVEFRONT): a pass holds
// whole samples of every pixel.  Primary rays have no jitter (raytraceBVH.comp:338), so one
// primary ray per pixel is traced, one thread per ray; k_wave_first then runs each pixel's
// samples in order (carry, RNG state, first shading off the pixel's primary surface) and
// queues the paths that scatter, those off dielectrics in a queue of their own;
// k_wave_finish carries every queued path to its end in one persistent launch, the
// dielectric queue first (those paths enter closed meshes and bounce longest); the samples
// are added to the image in order.  Path state lives in HBM between launches.
//
// One persistent launch (smaller images, RT_RENDER_PERSISTENT_PATHS and the layout flags):
// lanes take work items from 8x8 pixel tiles and run their samples to completion, advancing
// every active lane by one traversal step per iteration.  A pixel job keeps the accumulator
// in registers and traces the shared first segment once per pixel.  A split job gives each
// (pixel, group of samples) its own lane: the carry entering sample s is a pure function of
// the image alpha, the pixel's seed and s (each sample's first draw is the next carry), so a
// lane walks the chain to its sample, stores its colours, and a second pass adds the samples
// in order.  One launch may render several images (jobs), each with its own scene, camera,
// image and sample count: the work items of all jobs are concatenated, each job starting on a
// chunk boundary, so a wave's chunk belongs to one job and is found with one wave-uniform
// search of the job table.  Rendering spp = a and then spp = b into the same image equals
// rendering spp = a + b, which lets a scheduler hand out samples in slices.
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
    const uint32_t *gate;  // k_sum_samples: when set, sum only if a finish launch took all its queue (pass_complete)
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

// Wave-level choice between shading and tracing.  Shading (RNG draws, sampling, scatter, the
// next segment's setup) is long and only a few lanes finish a segment per iteration, so a wave
// shades when shade_min of its lanes hold a resolved segment -- or a third of its live lanes,
// when fewer are live -- or when none is left tracing; the resolved lanes wait meanwhile.
// Without the second bound a draining wave would shade only once every lane had resolved, and
// its last paths would advance in lockstep, each bounce waiting for the slowest lane.
__device__ __forceinline__ bool shade_now(unsigned long long mres, unsigned long long mtr, uint32_t shade_min) {
    if (mres == 0ull) return false;
    if (mtr == 0ull) return true;
    const uint32_t live = (uint32_t)__popcll(mres | mtr);
    return (uint32_t)__popcll(mres) >= min(shade_min, max(1u, live / 3u));
}

// The surface a segment ended on (raytraceBVH.comp:280-300): hit point, face-forward normal
// and material.  It depends on the segment and its hit only, so the samples of a pixel that
// share a primary hit share it.
struct Surface {
    v3 p, normal, alb;
    float w;        // albedo.w: metal fuzz, dielectric index of refraction
    uint32_t type;  // materialType
    int bf;         // back face
    bool hit;
};

__device__ __forceinline__ Surface surface_at(const RenderJob &J, const LeafRecord *rec, uint32_t hslot, float ht,
                                              v3 o, v3 d) {
    Surface s{};
    s.hit = hslot != kEmpty;
    if (!s.hit) return s;
    const LeafRecord r = rec[hslot];
    const rt_material m = J.mats[J.slot_mat[hslot]];
    s.alb = v3{m.albedo[0], m.albedo[1], m.albedo[2]};
    s.w = m.albedo[3];
    s.type = m.materialType;
    s.p = o + ht * d;
    v3 n;
    if (is_sphere_record(r)) n = (s.p - v3{r.r0.x, r.r0.y, r.r0.z}) / r.r0.w;
    else n = v3{r.r1.w, r.r2.w, r.r3.w};
    s.bf = dot(d, n) > 0.0f ? 1 : 0;
    s.normal = (float)(1 - 2 * s.bf) * n;
    return s;
}

// Adds the surface's emitted light and scatters the incoming direction d
// (raytraceBVH.comp:94-115, 276-308).  Returns true when the path continues, with (o, d)
// replaced by the next segment's ray; the caller owns the bounce count.  A miss adds the
// background.
__device__ __forceinline__ bool scatter_at(const RenderJob &J, const Surface &s, v3 &o, v3 &d, uint32_t &st,
                                           v3 &color, v3 &atten) {
    if (!s.hit) {
        const v3 background{0.0f, 0.0f, 0.0f};  // _BACKGROUND_COLOR
        color = color + mul(background, atten);
        return false;
    }
    const v3 emitted = s.type == RT_LIGHT ? s.alb : v3{0.0f, 0.0f, 0.0f};
    color = color + mul(emitted, atten);
    bool scattered = false;
    v3 att{0.0f, 0.0f, 0.0f}, nd{0.0f, 0.0f, 0.0f};
    const bool ref_mats = (J.flags & RT_RENDER_REFERENCE_MATERIALS) != 0;
    if (s.type == RT_DIFFUSE || (s.type == RT_METALLIC && !ref_mats)) {
        // both draw one point in the unit sphere: one pass of the sampling code for a wave
        // that holds both kinds of lanes
        att = s.alb;
        const v3 rs = random_in_unit_sphere(st);
        if (s.type == RT_DIFFUSE) {
            nd = normalize(s.normal + normalize(rs));
            scattered = true;
        } else {
            nd = normalize(reflect(d, s.normal) + s.w * rs);
            scattered = dot(nd, s.normal) > 0.0f;
        }
    } else if (s.type == RT_DIELECTRIC && !ref_mats) {
        att = v3{1.0f, 1.0f, 1.0f};
        const float ri = s.bf ? s.w : (1.0f / s.w);
        const float ct = gmin(dot(-d, s.normal), 1.0f);
        const float stt = sqrtf(1.0f - ct * ct);
        const bool cannot = ri * stt > 1.0f;
        if (cannot || schlick(ct, ri) > rng_next(st)) nd = reflect(d, s.normal);
        else nd = refract(d, s.normal, ri);
        nd = normalize(nd);
        scattered = true;
    }
    if (!scattered) return false;
    atten = mul(atten, att);
    o = s.p;
    d = nd;
    return true;
}

// One resolved segment of a path: its surface, then the scatter.  hslot == kEmpty is a miss.
__device__ __forceinline__ bool shade_segment(const RenderJob &J, const LeafRecord *rec, uint32_t hslot, float ht,
                                              v3 &o, v3 &d, uint32_t &st, v3 &color, v3 &atten) {
    return scatter_at(J, surface_at(J, rec, hslot, ht, o, d), o, d, st, color, atten);
}

// a resolved segment of a persistent-kernel path
template <bool MULTI>
__device__ __forceinline__ void shade(const RenderArgs &a, PixelState &ps, Trav &t, uint32_t &segs) {
    const RenderJob &J = job_of<MULTI>(a, ps);
    if (ps.bounce == 0 && !ps.pc_valid && !(J.flags & RT_RENDER_NO_PRIMARY_CACHE)) {
        ps.pc_slot = ps.hslot;
        ps.pc_t = ps.ht;
        ps.pc_valid = true;
    }
    if (!shade_segment(J, rec_of<MULTI>(a, ps), ps.hslot, ps.ht, ps.o, ps.d, ps.st, ps.color, ps.atten)) {
        end_sample<MULTI>(a, ps, t, segs);
        return;
    }
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
        // A lane's own sequence of segments is the same whichever phase the wave picks.
        const unsigned long long mres = __ballot(ps.phase == kResolved), mtr = __ballot(ps.phase == kTracing);
        if (shade_now(mres, mtr, a.shade_min)) {
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

// A bounce-by-bounce pass's counters: its two queues' lengths, the finish kernel's work counter
// and, when that launch yielded, the counter's value at the yield (see k_wave_finish).  The pass
// is complete once every queued path was taken before any yield.
constexpr uint32_t kYieldBit = 0x80000000u;  // set in the work counter by a launch that yields

__device__ __forceinline__ bool pass_complete(const uint32_t *count) {
    return ((count[2] & kYieldBit) ? count[3] : count[2]) >= count[0] + count[1];
}

template <bool MULTI> __global__ void k_sum_samples(RenderArgs a) {
    if (a.gate && !pass_complete(a.gate)) return;
    if constexpr (MULTI) sum_job(a.jobs[blockIdx.y]);
    else sum_job(a.job0);
}

// ---------------------------------------------------------------------------
// Bounce-by-bounce schedule.  The paths of a pass are its (sample, pixel) pairs, numbered
// sample-major (p = k * npix + pix), which is also the layout of the pass's sample storage
// that k_sum_samples adds in order.  Per path the storage holds {colour, carry out} and a
// second array {attenuation, RNG state}; a queue entry is the segment's ray and its path.
// Every path performs the persistent kernel's operations in the same order, so images are
// identical; only the schedule differs.
// ---------------------------------------------------------------------------
struct WaveArgs {
    RenderJob J;
    uint32_t npix, nsamp;     // pixels; samples of this pass (paths = npix * nsamp)
    uint32_t prim_per_pixel;  // the first bounce's hits are per pixel (primary cache), else per path
    uint32_t bounce;          // bounce being shaded (0: primary segments)
    const rt_ray *rays_in;
    const uint32_t *ids_in;
    const rt_hit *hits_in;
    const uint32_t *n_in;
    rt_ray *rays_out;
    uint32_t *ids_out;
    uint32_t *n_out;
    // k_wave_first only: when set, paths that scatter off a dielectric go to this queue instead,
    // and the finish kernel takes them first (rays_in2 ...): they are the ones that enter the
    // closed meshes and bounce longest, so starting them early keeps them out of the drain.
    rt_ray *rays_out2;
    uint32_t *ids_out2;
    uint32_t *n_out2;
    const rt_ray *rays_in2;
    const uint32_t *ids_in2;
    const uint32_t *n_in2;
    float4 *col;  // [path] {colour, carry out}
    float4 *att;  // [path] {attenuation, RNG state bits}
    unsigned long long *segments;
};

__device__ __forceinline__ void store_ray(rt_ray *rays, uint32_t i, v3 o, v3 d) {
    float4 *q = reinterpret_cast<float4 *>(rays + i);
    q[0] = make_float4(o.x, o.y, o.z, 0.001f);        // sceneHit's tMin (raytraceBVH.comp:268)
    q[1] = make_float4(d.x, d.y, d.z, 10000000.0f);   // and tMax (:269)
}

// primary rays: one per pixel (prim_per_pixel) or one per path, in path order
__global__ void k_wave_primary(WaveArgs a) {
    const uint32_t n = a.prim_per_pixel ? a.npix : a.npix * a.nsamp;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t pix = i % a.npix;
        v3 o, d;
        primary(a.J, pix % a.J.width, pix / a.J.width, o, d);
        store_ray(a.rays_out, i, o, d);
    }
    if (a.segments && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(a.segments, (unsigned long long)n);
}

constexpr uint32_t kShadeBlock = 512;

// Appends the block's continuing paths to the queue with one atomic per queue: the block's
// waves add their counts in LDS first (a counter hit by every wave would serialise them).
// Each thread offers G paths at once (p0, p0 + pstride, ...), so a pass's samples take G times
// fewer block-wide rounds.  Paths with `front` set go to the second queue when the launch has
// one.  Every thread of the block calls it.
struct BlockAppend {
    uint32_t wave_off[2][kShadeBlock / 64];
    uint32_t base[2];
};

template <int G>
__device__ __forceinline__ void append_block(BlockAppend &sh, const WaveArgs &a, const bool (&cont)[G],
                                             const bool (&front)[G], const v3 (&o)[G], const v3 (&d)[G],
                                             uint32_t p0, uint32_t pstride) {
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6;
    const unsigned long long lt = lane ? (~0ull >> (64u - lane)) : 0ull;
    const bool two = a.n_out2 != nullptr;
    unsigned long long live[2][G];
    uint32_t cnt[2] = {0u, 0u};
#pragma unroll
    for (int j = 0; j < G; ++j) {
        const bool q = front[j] && two;
        live[0][j] = __ballot(cont[j] && !q);
        live[1][j] = __ballot(cont[j] && q);
        cnt[0] += (uint32_t)__popcll(live[0][j]);
        cnt[1] += (uint32_t)__popcll(live[1][j]);
    }
    if (lane == 0) {
        sh.wave_off[0][wid] = cnt[0];
        sh.wave_off[1][wid] = cnt[1];
    }
    __syncthreads();
    if (threadIdx.x < 2u) {  // exclusive offsets of the block's waves, one global append per queue
        const uint32_t k = threadIdx.x;
        uint32_t total = 0;
        for (uint32_t w = 0; w < blockDim.x / 64u; ++w) {
            const uint32_t c = sh.wave_off[k][w];
            sh.wave_off[k][w] = total;
            total += c;
        }
        sh.base[k] = total ? atomicAdd(k ? a.n_out2 : a.n_out, total) : 0u;
        if (total && a.segments) atomicAdd(a.segments, (unsigned long long)total);
    }
    __syncthreads();
    uint32_t before[2] = {0u, 0u};  // this wave's paths of the earlier samples of the group
#pragma unroll
    for (int j = 0; j < G; ++j) {
        const uint32_t q = (front[j] && two) ? 1u : 0u;
        if (cont[j]) {
            const uint32_t slot =
                sh.base[q] + sh.wave_off[q][wid] + before[q] + (uint32_t)__popcll(live[q][j] & lt);
            store_ray(q ? a.rays_out2 : a.rays_out, slot, o[j], d[j]);
            (q ? a.ids_out2 : a.ids_out)[slot] = p0 + (uint32_t)j * pstride;
        }
        before[0] += (uint32_t)__popcll(live[0][j]);
        before[1] += (uint32_t)__popcll(live[1][j]);
    }
    __syncthreads();  // the offsets are rewritten by the next call
}

constexpr int kFirstGroup = 4;  // samples a bounce-0 thread appends at once

// Bounce 0 of a pass, one thread per pixel: its samples in order, each starting from the carry
// the previous one left (raytraceBVH.comp:353-375), from the pixel's primary ray and its
// primary hit; the paths that scatter are queued.
__global__ __launch_bounds__(kShadeBlock) void k_wave_first(WaveArgs a) {
    __shared__ BlockAppend sh;
    const RenderJob &J = a.J;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (a.npix + stride - 1) / stride;  // the same for every thread of the grid
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t pix = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        const bool live = pix < a.npix;
        uint32_t x = 0, y = 0, seed0 = 0;
        float carry = 0.0f;
        v3 o0{0.0f, 0.0f, 0.0f}, d0{0.0f, 0.0f, 0.0f};
        Surface surf{};  // the pixel's primary hit surface, shared by its samples
        if (live) {
            x = pix % J.width;
            y = pix / J.width;
            seed0 = rng_seed(x, y, J.random_state);
            carry = reinterpret_cast<const float4 *>(J.accum)[pix].w;
            primary(J, x, y, o0, d0);
            if (a.prim_per_pixel && J.depth > 0) {
                const rt_hit h = a.hits_in[pix];
                surf = surface_at(J, J.rec, h.prim, h.t, o0, d0);
            }
        }
        for (uint32_t k0 = 0; k0 < a.nsamp; k0 += kFirstGroup) {
            bool cont[kFirstGroup], front[kFirstGroup];
            v3 o[kFirstGroup], d[kFirstGroup];
#pragma unroll
            for (int j = 0; j < kFirstGroup; ++j) {
                const uint32_t k = k0 + (uint32_t)j;
                const uint32_t p = k * a.npix + pix;
                cont[j] = false;
                o[j] = o0;
                d[j] = d0;
                if (live && k < a.nsamp) {
                    uint32_t st = seed0 + f2u_sat(carry * 4294967294.0f);  // the literal is 2^32 in fp32
                    const float next = rng_next(st);
                    carry = next;
                    v3 color{0.0f, 0.0f, 0.0f}, atten{1.0f, 1.0f, 1.0f};
                    if (!a.prim_per_pixel && J.depth > 0) {
                        const rt_hit h = a.hits_in[p];
                        surf = surface_at(J, J.rec, h.prim, h.t, o[j], d[j]);
                    }
                    cont[j] = scatter_at(J, surf, o[j], d[j], st, color, atten) && 1u < J.depth;
                    a.col[p] = make_float4(color.x, color.y, color.z, next);
                    if (cont[j]) a.att[p] = make_float4(atten.x, atten.y, atten.z, __uint_as_float(st));
                }
                front[j] = surf.type == RT_DIELECTRIC;
            }
            append_block<kFirstGroup>(sh, a, cont, front, o, d, k0 * a.npix + pix, a.npix);
        }
    }
}

// Shades the queued segments of one bounce (a.bounce) and queues the continuing paths.
__global__ __launch_bounds__(kShadeBlock) void k_wave_shade(WaveArgs a) {
    __shared__ BlockAppend sh;
    const RenderJob &J = a.J;
    const uint32_t n = *a.n_in;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (n + stride - 1) / stride;  // the same for every thread of the grid
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint32_t i = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool cont = false;
        uint32_t p = 0;
        v3 o{0.0f, 0.0f, 0.0f}, d{0.0f, 0.0f, 0.0f};
        if (i < n) {
            p = a.ids_in[i];
            const float4 *q = reinterpret_cast<const float4 *>(a.rays_in + i);
            const float4 r0 = q[0], r1 = q[1];
            o = v3{r0.x, r0.y, r0.z};
            d = v3{r1.x, r1.y, r1.z};
            const float4 c = a.col[p], t = a.att[p];
            v3 color{c.x, c.y, c.z}, atten{t.x, t.y, t.z};
            uint32_t st = __float_as_uint(t.w);
            const rt_hit h = a.hits_in[i];
            cont = shade_segment(J, J.rec, h.prim, h.t, o, d, st, color, atten) && a.bounce + 1 < J.depth;
            a.col[p] = make_float4(color.x, color.y, color.z, c.w);
            if (cont) a.att[p] = make_float4(atten.x, atten.y, atten.z, __uint_as_float(st));
        }
        const bool c1[1] = {cont}, f1[1] = {false};
        const v3 o1[1] = {o}, d1[1] = {d};
        append_block<1>(sh, a, c1, f1, o1, d1, p, 0u);
    }
}

// Runs the paths of a queue to their ends in one persistent launch.  The late bounces hold
// a small share of the paths, mostly rays that leave a surface inside a closed mesh and are
// long to trace; bounce-by-bounce launches would each wait for their longest ray, while
// here a lane continues its own path as soon as its segment is shaded.  Lanes take queue
// entries like the ray-query kernel takes rays, and shade in phases like k_render.
//
// Yielding: with `yield` set (a counter of yield requests in host memory), the wave that takes
// every kYieldPoll-th chunk of entries reads it; once it reads `yield_at` or more, it sets
// kYieldBit in the work counter and records the counter's value before that, work[1].  Every
// chunk below that value was taken before the bit and is run to its end; a wave whose take
// returns the bit drops that chunk and takes no more, so the launch ends with the drain of the
// paths in flight.  The entries left are exactly [work[1], n): k_resume_counter puts the counter
// back there, and a launch with the same arguments runs them.  Every path still performs the same
// operations in the same order, so the image does not change.  A server uses this to hand the
// device to short requests without waiting for the whole launch.  Reads of host memory cross the
// host link, so only one take in kYieldPoll reads it; the others cost nothing extra.
constexpr uint32_t kYieldPoll = 32;

struct FinishArgs {
    WaveArgs w;
    unsigned int *work;  // work[0]: the counter; work[1]: its value when the launch yielded
    uint32_t refill, leaf_weight, shade_min;
    uint32_t *yield;  // host-written yield requests (fine-grained host memory), or nullptr
    uint32_t yield_at;
};

template <int W, int S> __global__ __launch_bounds__(256) void k_wave_finish(FinishArgs f) {
    extern __shared__ uint32_t lds[];
    uint2 spill[kSpillCap];
    const StackMem m = stack_mem<S>(lds, spill);
    const WaveArgs &a = f.w;
    const RenderJob &J = a.J;
    const uint32_t n2 = a.n_in2 ? *a.n_in2 : 0u;  // the second queue's entries come first
    const uint32_t n = *a.n_in + n2;
    const uint32_t lane = threadIdx.x & 63u;
    const unsigned long long lt = lane ? (~0ull >> (64u - lane)) : 0ull;
    const uint32_t chunk = kRenderChunk;  // queue entries per atomic
    uint32_t phase = kIdle, p = 0, bounce = 0, st = 0, hslot = kEmpty;
    float next = 0.0f, ht = 0.0f;
    v3 color{0.0f, 0.0f, 0.0f}, atten{1.0f, 1.0f, 1.0f};
    Trav t;
    t.have = false;
    t.top = 0;
    uint32_t segs = 0;  // continuation segments (the queue's own were counted when queued)
    bool exhausted = false;
    uint32_t q_next = 0, q_end = 0;
    for (;;) {
        const unsigned long long idle = __ballot(phase == kIdle);
        const uint32_t nidle = (uint32_t)__popcll(idle);
        if (!exhausted && nidle >= f.refill) {
            if (q_next == q_end) {
                const int leader = __ffsll((unsigned long long)idle) - 1;
                uint32_t base = 0;
                if ((int)lane == leader) {
                    base = atomicAdd(f.work, chunk);
                    if (f.yield && !(base & kYieldBit) && (base / chunk) % kYieldPoll == 0u &&
                        __hip_atomic_load(f.yield, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) >= f.yield_at) {
                        const uint32_t before = atomicOr(f.work, kYieldBit);
                        if (!(before & kYieldBit)) f.work[1] = before;  // the first to yield records where
                    }
                }
                base = (uint32_t)__shfl((int)base, leader);
                if (base >= n) exhausted = true;  // taken after a yield (kYieldBit), or past the end
                else { q_next = base; q_end = min(base + chunk, n); }
            }
            if (!exhausted) {
                const uint32_t take = min(nidle, q_end - q_next);
                const uint32_t r = (uint32_t)__popcll(idle & lt);
                if (phase == kIdle && r < take) {
                    const uint32_t i = q_next + r;
                    const bool second = i < n2;
                    const uint32_t qi = second ? i : i - n2;
                    p = (second ? a.ids_in2 : a.ids_in)[qi];
                    const float4 *q = reinterpret_cast<const float4 *>((second ? a.rays_in2 : a.rays_in) + qi);
                    const float4 r0 = q[0], r1 = q[1];
                    const float4 c = a.col[p], at = a.att[p];
                    color = v3{c.x, c.y, c.z};
                    next = c.w;
                    atten = v3{at.x, at.y, at.z};
                    st = __float_as_uint(at.w);
                    bounce = a.bounce;
                    trav_init(t, v3{r0.x, r0.y, r0.z}, v3{r1.x, r1.y, r1.z}, r0.w, r1.w, J.nwide);
                    phase = kTracing;
                }
                q_next += take;
            }
        }
        if (__ballot(phase != kIdle) == 0ull) break;
        const unsigned long long mres = __ballot(phase == kResolved), mtr = __ballot(phase == kTracing);
        if (shade_now(mres, mtr, f.shade_min)) {
            if (phase == kResolved) {
                v3 o = t.o, d = t.d;
                if (shade_segment(J, J.rec, hslot, ht, o, d, st, color, atten) && bounce + 1 < J.depth) {
                    bounce++;
                    trav_init(t, o, d, 0.001f, 10000000.0f, J.nwide);
                    phase = kTracing;
                    segs++;
                } else {
                    a.col[p] = make_float4(color.x, color.y, color.z, next);
                    phase = kIdle;
                }
            }
        } else {
            const bool tracing = phase == kTracing;
            const bool on_node = tracing && at_node(t), on_leaf = tracing && at_leaf(t);
            if (choose_leaf_phase(on_node, on_leaf, f.leaf_weight)) {
                if (on_leaf) leaf_step<S, false, false>(t, m, J.rec);
            } else {
                if (on_node) node_step<W, S, false, false>(t, m, J.wide);
            }
            if (tracing && trav_done(t)) {
                hslot = t.bslot;
                ht = t.bt;
                phase = kResolved;
            }
        }
    }
    if (a.segments) {
        unsigned long long sg = segs;
        for (int off = 32; off > 0; off >>= 1) sg += __shfl_xor(sg, off);
        if (lane == 0 && sg) atomicAdd(a.segments, sg);
    }
}

template <int W> void *pick_finish_s(uint32_t s) {
    return s == 4 ? (void *)k_wave_finish<W, 4> : s == 16 ? (void *)k_wave_finish<W, 16> : (void *)k_wave_finish<W, 8>;
}

void *pick_finish(uint32_t width, uint32_t s) {
    switch (width) {
        case 2: return pick_finish_s<2>(s);
        case 8: return pick_finish_s<8>(s);
        default: return pick_finish_s<4>(s);
    }
}

__global__ void k_accum_clear(float4 *accum, uint32_t npix) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npix; i += gridDim.x * blockDim.x)
        accum[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
}

// SingleTriangleFullScreen.frag:13-21
__global__ void k_resolve(const float4 *accum, uint32_t npix, uint32_t spp, float4 *out, const uint32_t *gate) {
    if (gate && !pass_complete(gate)) return;
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

// Before a yielded finish launch is resumed: its work counter back to where the yield left it.
__global__ void k_resume_counter(unsigned int *work) {
    if (threadIdx.x == 0 && (work[0] & kYieldBit)) work[0] = work[1];
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

void launch_resolve(const float *accum, uint32_t npix, uint32_t spp, float *out, hipStream_t stream,
                    const uint32_t *gate) {
    if (npix == 0) return;
    RT_REQUIRE(accum && out, "accum/out is NULL");
    uint32_t grid = std::min<uint32_t>((npix + 255) / 256, 65536);
    k_resolve<<<grid, 256, 0, stream>>>(reinterpret_cast<const float4 *>(accum), npix, spp,
                                        reinterpret_cast<float4 *>(out), gate);
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

namespace {

uint64_t wave_paths_per_pass() {
    uint64_t v = 1ull << 25;
    if (const char *e = std::getenv("RTGPU_WAVE_PATHS")) v = (uint64_t)std::max(1, std::atoi(e));
    return v;
}

}  // namespace

// per path: {colour, carry}, {attenuation, state}, two queues of (ray, path), hits
uint64_t wave_scratch_bytes(uint64_t paths) { return paths * (16 + 16 + 2 * (32 + 4) + 16) + 256; }

uint64_t per_pass_limit(uint64_t npix) { return std::max<uint64_t>(1, wave_paths_per_pass() / npix); }

namespace {
// First bounce whose segments k_wave_finish traces; the earlier ones go bounce by bounce.
// Bounce 0 (primary segments, traced once per pixel, and every path's first shading) is the
// bulk of the work and fills the device; from bounce 1 on, paths thin out and lengthen.
uint32_t wave_finish_bounce() {
    uint32_t v = 1;
    if (const char *e = std::getenv("RTGPU_WAVE_FINISH")) v = (uint32_t)std::max(1, std::atoi(e));
    return v;
}
}  // namespace

// A yieldable pass: its finish launch and gated sum, launched again until the pass is complete.
struct WaveResume {
    FinishArgs fa;
    void *kernel;
    uint32_t grid, block;
    size_t lds;
    RenderArgs sum;
    uint32_t npix;
    uint32_t *count;  // the pass's queue lengths, work counter and the counter's value at a yield
};

namespace {
void launch_finish(const WaveResume &r, hipStream_t stream) {
    FinishArgs fa = r.fa;
    void *args[] = {&fa};
    RT_HIP(hipLaunchKernel(r.kernel, dim3(r.grid), dim3(r.block), args, r.lds, stream));
    RT_HIP(hipGetLastError());
}
}  // namespace

void resume_render_wave(const WaveResume &r, const WaveYield &yield, hipStream_t stream) {
    WaveResume again = r;
    again.fa.yield = yield.flag;
    again.fa.yield_at = yield.yield_at;
    k_resume_counter<<<1, 64, 0, stream>>>(again.fa.work);
    RT_HIP(hipGetLastError());
    launch_finish(again, stream);
    RenderArgs sa = r.sum;
    launch_sum_samples(sa, false, r.npix, stream);
}

const uint32_t *wave_counters(const WaveResume &r) { return r.count; }

bool wave_pass_complete(const uint32_t *c) { return ((c[2] & kYieldBit) ? c[3] : c[2]) >= c[0] + c[1]; }

// Bounce-by-bounce rendering of one image: passes of whole samples of every pixel, each
// continuing the carry chain of the previous one.  A pass holds at most wave_paths_per_pass()
// paths, and no more than `scratch` holds when the caller provides it (else the storage comes
// from the context's scratch cache).
void launch_render_wave(rt_context *ctx, const Scene &sc, const RenderJob &j, unsigned long long *segments,
                        const LaunchCfg &c, hipStream_t stream, char *scratch, uint64_t scratch_bytes,
                        rt_hit *primary, bool primary_ready, const WaveYield *yield,
                        std::shared_ptr<WaveResume> *resume) {
    const uint32_t npix = j.width * j.height;
    uint64_t per_pass = std::min<uint64_t>(j.spp, per_pass_limit(npix));
    if (scratch) {
        RT_REQUIRE(scratch_bytes >= wave_scratch_bytes(npix), "render scratch smaller than one sample of the image");
        per_pass = std::min<uint64_t>(per_pass, std::max<uint64_t>(1, (scratch_bytes - 256) / (wave_scratch_bytes(npix) - 256)));
    }
    if (resume) resume->reset();
    // only one pass whose last bounces run in the finish kernel, with the dielectric queue, can yield
    if (yield && !(resume && per_pass >= j.spp && j.depth > 1 && wave_finish_bounce() == 1)) yield = nullptr;
    const uint64_t maxp = (uint64_t)npix * per_pass;
    RT_REQUIRE(maxp < (1ull << 32), "too many paths in one pass");
    LaunchCfg tc = c;
    tc.persistent = 1;
    std::unique_ptr<ScratchLease> lease;
    if (!scratch) lease = std::make_unique<ScratchLease>(ctx->scratch, wave_scratch_bytes(maxp), stream);
    char *mem = scratch ? scratch : lease->ptr();
    float4 *col = reinterpret_cast<float4 *>(mem);
    float4 *att = col + maxp;
    rt_ray *rays[2] = {reinterpret_cast<rt_ray *>(att + maxp), nullptr};
    rays[1] = rays[0] + maxp;
    rt_hit *hits = reinterpret_cast<rt_hit *>(rays[1] + maxp);
    uint32_t *ids[2] = {reinterpret_cast<uint32_t *>(hits + maxp), nullptr};
    ids[1] = ids[0] + maxp;
    uint32_t *count = ids[1] + maxp;  // two queue counters, a work counter and its value at a yield
    const bool cache = !(j.flags & RT_RENDER_NO_PRIMARY_CACHE);
    if (!cache) primary = nullptr;
    const uint32_t block = kShadeBlock;
    // the shading kernels wait on scattered path state and leaf records: many waves per CU
    const uint32_t full = (uint32_t)ctx->num_cus * 8u;
    for (uint32_t s0 = 0; s0 < j.spp; s0 += (uint32_t)per_pass) {
        const uint32_t nsamp = (uint32_t)std::min<uint64_t>(per_pass, j.spp - s0);
        const uint64_t paths = (uint64_t)npix * nsamp;
        WaveArgs a{};
        a.J = j;
        a.npix = npix;
        a.nsamp = nsamp;
        a.prim_per_pixel = cache ? 1u : 0u;
        a.col = col;
        a.att = att;
        a.segments = segments;
        const auto grid_for = [&](uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + block - 1) / block); };
        rt_hit *first_hits = primary ? primary : hits;
        if (j.depth > 0 && !(primary && primary_ready)) {
            const uint64_t nprim = cache ? npix : paths;
            a.rays_out = rays[0];
            k_wave_primary<<<grid_for(nprim), block, 0, stream>>>(a);
            RT_HIP(hipGetLastError());
            launch_trace_queue(ctx, sc, rays[0], (uint32_t)nprim, nullptr, first_hits, tc, stream);
            primary_ready = primary != nullptr;
        }
        int cur = 0;
        const uint32_t finish_at = wave_finish_bounce();
        k_zero<<<1, 64, 0, stream>>>(count + cur);
        a.bounce = 0;
        a.hits_in = first_hits;
        a.rays_out = rays[cur];
        a.ids_out = ids[cur];
        a.n_out = count + cur;
        if (finish_at == 1) {  // dielectric scatters into the second queue, which the finish kernel takes first
            k_zero<<<1, 64, 0, stream>>>(count + 1);
            a.rays_out2 = rays[1];
            a.ids_out2 = ids[1];
            a.n_out2 = count + 1;
        }
        k_wave_first<<<std::min(full, grid_for(npix)), block, 0, stream>>>(a);
        RT_HIP(hipGetLastError());
        a.rays_in2 = a.rays_out2;
        a.ids_in2 = a.ids_out2;
        a.n_in2 = a.n_out2;
        a.rays_out2 = nullptr;
        a.ids_out2 = nullptr;
        a.n_out2 = nullptr;
        for (uint32_t b = 1; b < j.depth; ++b) {
            if (b >= finish_at) {  // the remaining bounces of the queued paths in one launch
                a.bounce = b;
                a.rays_in = rays[cur];
                a.ids_in = ids[cur];
                a.n_in = count + cur;
                WaveResume w{};
                w.fa = FinishArgs{a, count + 2, tc.refill, tc.leaf_weight, tc.shade_min,
                                  yield ? yield->flag : nullptr, yield ? yield->yield_at : 0u};
                k_zero<<<1, 64, 0, stream>>>(w.fa.work);
                w.lds = stack_lds_bytes(tc.lds_stack, tc.block);
                w.kernel = pick_finish(sc.width, tc.lds_stack);
                int per_cu = 0;
                RT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, w.kernel, (int)tc.block, w.lds));
                w.grid = (uint32_t)std::max(1, per_cu) * (uint32_t)ctx->num_cus;
                w.block = tc.block;
                launch_finish(w, stream);
                if (yield) {
                    w.npix = npix;
                    w.count = count;
                    *resume = std::make_shared<WaveResume>(w);
                }
                break;
            }
            launch_trace_queue(ctx, sc, rays[cur], (uint32_t)paths, count + cur, hits, tc, stream);
            k_zero<<<1, 64, 0, stream>>>(count + (cur ^ 1));
            a.bounce = b;
            a.rays_in = rays[cur];
            a.ids_in = ids[cur];
            a.hits_in = hits;
            a.n_in = count + cur;
            a.rays_out = rays[cur ^ 1];
            a.ids_out = ids[cur ^ 1];
            a.n_out = count + (cur ^ 1);
            k_wave_shade<<<std::min(full, grid_for(paths)), block, 0, stream>>>(a);
            RT_HIP(hipGetLastError());
            cur ^= 1;
        }
        RenderArgs sa{};  // the pass's samples, added to the image in order
        sa.job0 = j;
        sa.job0.group = 1;
        sa.job0.samples = col;
        sa.job0.spp = nsamp;
        if (yield) {  // a yielded pass is summed by the launch that completes it
            sa.gate = count;
            (*resume)->sum = sa;
        }
        launch_sum_samples(sa, false, npix, stream);
    }
}

bool wave_schedule(uint32_t width, uint32_t height, uint32_t spp, uint32_t flags) {
    const bool layout = (flags & (RT_RENDER_SPLIT_SAMPLES | RT_RENDER_NO_SPLIT | RT_RENDER_PERSISTENT_PATHS)) != 0;
    RT_REQUIRE(!(layout && (flags & RT_RENDER_WAVEFRONT)), "RT_RENDER_WAVEFRONT excludes the persistent-launch flags");
    return (flags & RT_RENDER_WAVEFRONT) || (!layout && (uint64_t)width * height * spp >= kWaveMinSamples);
}

void launch_render(rt_context *ctx, const Scene &sc, const rt_params &p, uint32_t width, uint32_t height, uint32_t spp,
                   uint32_t flags, float *accum, unsigned long long *segments, const LaunchCfg &c,
                   hipStream_t stream, char *scratch, uint64_t scratch_bytes) {
    RenderJob j = make_render_job(sc, p, width, height, spp, flags, accum);
    if (spp == 0) return;
    check_stack(sc, c);
    if (wave_schedule(width, height, spp, flags)) {
        if (scratch && scratch_bytes < wave_scratch_bytes((uint64_t)width * height))
            throw Error(RT_ERR_BUDGET, "one sample of a " + std::to_string(width) + "x" + std::to_string(height) +
                                           " image needs " + std::to_string(wave_scratch_bytes((uint64_t)width * height)) +
                                           " bytes of path storage; the scratch holds " + std::to_string(scratch_bytes));
        launch_render_wave(ctx, sc, j, segments, c, stream, scratch, scratch_bytes);
        return;
    }
    float4 *samples = nullptr;
    uint32_t group = split_group(ctx, width, height, spp, flags);
    if (scratch && group && 16ull * width * height * spp > scratch_bytes) {
        // the per-sample storage does not fit: whole pixels per lane give the same image
        if (flags & RT_RENDER_SPLIT_SAMPLES)
            throw Error(RT_ERR_BUDGET, "RT_RENDER_SPLIT_SAMPLES needs " + std::to_string(16ull * width * height * spp) +
                                           " bytes of sample storage; the scratch holds " +
                                           std::to_string(scratch_bytes));
        group = 0;
    }
    const bool split = group != 0;
    std::unique_ptr<ScratchLease> lease;  // the samples' storage, back in the cache after the sum
    if (split) {
        if (!scratch) lease = std::make_unique<ScratchLease>(ctx->scratch, 16ull * width * height * spp, stream);
        samples = reinterpret_cast<float4 *>(scratch ? scratch : lease->ptr());
        make_split(j, samples, group);
    }
    RenderArgs a{};
    a.job0 = j;
    a.jobs = nullptr;
    a.njobs = 1;
    a.nitems = j.nitems;
    a.segments = segments;
    launch_render_kernel(ctx, a, sc.width, false, c, stream);
    if (split) launch_sum_samples(a, false, width * height, stream);
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
