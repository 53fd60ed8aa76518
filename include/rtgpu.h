/*
 * rtgpu.h -- C ABI of the MI355X ray-tracing hot path.
 *
 * This is the drop-in boundary for the reference's Vulkan raytrace path.  Each
 * entry point names the reference interface it replaces.  Reference paths are
 * relative to RaytracerGPU_MastersProject/.
 *
 *   descriptor set of raytraceBVH.comp (RaytracerBVH.cpp:202-238, 677-685):
 *     binding 0 ParameterUBO      -> rt_params (byte-identical, 80 B)
 *     binding 1 RGBA32F image     -> float *accum (device, W*H*4, row-major)
 *     binding 2 Triangle SSBO     -> rt_triangle[] (64 B)
 *     binding 3 Sphere SSBO       -> rt_sphere[]   (32 B)
 *     binding 4 Material SSBO     -> rt_material[] (32 B)
 *     binding 5 HLBVHNode SSBO    -> rt_bvh_node[] (40 B, the RaytracerBVH builder output)
 *
 * Conventions: every entry point returns an rt_status (0 = RT_OK).  On failure
 * rt_last_error(ctx) holds a message; nothing is thrown across the ABI.  A
 * context is used from one host thread at a time.  "device" pointers are HIP
 * device allocations on the context's GPU; "host" pointers are borrowed for the
 * duration of the call only.  Streams are hipStream_t passed as void* (NULL =
 * the null stream); work is stream-ordered and the call returns without
 * waiting unless stated.
 */
#ifndef RTGPU_H
#define RTGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rt_status {
    RT_OK = 0,
    RT_ERR_INVALID = 1,    /* bad argument, size or layout                     */
    RT_ERR_NO_MEMORY = 2,  /* device allocation failed                         */
    RT_ERR_HIP = 3,        /* HIP runtime error                                */
    RT_ERR_NOT_FOUND = 4,  /* unknown scene / request id                       */
    RT_ERR_BUDGET = 5,     /* HBM budget of a server would be exceeded         */
    RT_ERR_OVERFLOW = 6,   /* traversal stack bound exceeded (reported, rays terminated) */
    RT_ERR_BUSY = 7        /* request not finished yet                         */
} rt_status;

typedef struct rt_context rt_context;

/* ---------------- byte layouts (see static_asserts in the implementation) ---------------- */

/* definitions.glsl:15-21 / SceneTypes.hpp:41-47 (GPU::Triangle), 64 B */
typedef struct rt_triangle {
    float v0[4], v1[4], v2[4];
    uint32_t materialIndex, modelIndex, _pad[2];
} rt_triangle;

/* definitions.glsl:23-28 / SceneTypes.hpp:62-66 (GPU::Sphere), 32 B */
typedef struct rt_sphere {
    float center[4];
    float radius;
    uint32_t materialIndex, modelIndex, _pad;
} rt_sphere;

/* definitions.glsl:10-13 / SceneTypes.hpp:82-84 (GPU::Material), 32 B.
 * albedo[3] is unused by the reference; here it carries fuzz (METALLIC) or the
 * index of refraction (DIELECTRIC). */
typedef struct rt_material {
    float albedo[4];
    uint32_t materialType; /* RT_LIGHT, RT_DIFFUSE, RT_METALLIC, RT_DIELECTRIC */
    uint32_t _pad[3];
} rt_material;

/* definitions.glsl:52-72 / SceneTypes.hpp:107-123 (GPU::BVHNode), 40 B.
 * 2N-1 nodes; 0 is the root; internal nodes 0..N-2; leaves N-1..2N-2;
 * a leaf has left == right == 0. */
typedef struct rt_bvh_node {
    float minX, maxX, minY, maxY, minZ, maxZ;
    uint32_t left, right, primitiveIndex, primitiveType;
} rt_bvh_node;

/* raytraceBVH.comp:7-18 (std140) / RaytracerBVH.hpp:31-42, 80 B */
typedef struct rt_params {
    float camPos[4], camLookAt[4], camUpDir[4];
    float verticalFOV; /* degrees */
    uint32_t numTriangles, numSpheres, numMaterials, numLights, maxRayTraceDepth, randomState;
    uint32_t _pad;
} rt_params;

/* ray query records: 32 B ray, 16 B hit (prim = RT_NO_HIT when nothing is hit; t, u, v: see RT_TRACE_FP32) */
typedef struct rt_ray { float o[3]; float tmin; float d[3]; float tmax; } rt_ray;
typedef struct rt_hit { uint32_t prim; float t; float u; float v; } rt_hit;

enum { RT_LIGHT = 0, RT_DIFFUSE = 1, RT_METALLIC = 2, RT_DIELECTRIC = 3 };   /* definitions.glsl:79-82 */
enum { RT_SPHERE_PRIMITIVE = 0, RT_TRIANGLE_PRIMITIVE = 1 };                   /* definitions.glsl:84-85 */
#define RT_NO_HIT 0xFFFFFFFFu
#define RT_SPHERE_BIT 0x80000000u /* rt_hit.prim of a sphere hit = RT_SPHERE_BIT | sphere index */

/* ---------------- context ---------------- */

/* Replaces Device{window} + Raytracer::initVulkan (RaytracerBVH.cpp:7-11, RaytracerBVH.hpp:145-178). */
int rt_create(rt_context **out, int device);
void rt_destroy(rt_context *ctx);
const char *rt_last_error(const rt_context *ctx);

/* ---------------- scenes ---------------- */

#define RT_BUILD_REFERENCE_LEAVES 1u /* ConstructHLBVH.comp:152-169: leaf N-1+i holds primitive i (unsorted) */
#define RT_BUILD_REFERENCE_MORTON 2u /* GenerateMortonCodesOfPrimitives.comp:60-65: coord/span quantisation */
#define RT_BUILD_KEEP_BINARY 4u      /* keep the 2N-1 binary nodes resident for rt_scene_get_nodes */

typedef struct rt_build_options {
    uint32_t flags;     /* RT_BUILD_* */
    uint32_t width;     /* wide-BVH width: 2, 4 or 8 (0 = 4) */
    uint32_t leaf_size; /* max primitives per wide leaf, 1..8 (0 = 1) */
    uint32_t _reserved;
} rt_build_options;

/* Upload a scene and build its BVH on the device.  Replaces RaytraceScene::prepForRender
 * (RaytraceScene.cpp:73-76, 157-200) plus compute stage S1 (RaytracerBVH.cpp:734-997):
 * model->world transform (when models != NULL; mat4 column-major, 16 floats each),
 * enclosing AABB, Morton codes, radix sort, Karras LBVH, refit; then the binary LBVH is
 * collapsed into the resident wide layout the kernels traverse.
 * nodes != NULL: use this reference builder output (2N-1 rt_bvh_node) instead of building.
 * Zero triangles and zero spheres is a valid (empty) scene. */
int rt_scene_upload(rt_context *ctx, uint32_t *scene_id, const rt_triangle *tris, uint32_t ntri,
                    const rt_sphere *spheres, uint32_t nsph, const rt_material *materials, uint32_t nmat,
                    const float *models, uint32_t nmodels, const rt_bvh_node *nodes, uint32_t nnodes,
                    const rt_build_options *opts);
int rt_scene_release(rt_context *ctx, uint32_t scene_id);

typedef struct rt_scene_info {
    uint32_t ntri, nsph, nmat, nprim;
    uint32_t width, leaf_size, nwide, wide_depth;
    uint32_t stack_bound, flags;
    uint64_t device_bytes; /* resident bytes (wide nodes + leaf records + tables) */
    float build_ms;        /* device time of the last build */
    float _pad;
} rt_scene_info;
int rt_scene_info_get(rt_context *ctx, uint32_t scene_id, rt_scene_info *info);

/* Download the RaytracerBVH builder output (requires RT_BUILD_KEEP_BINARY): 2N-1 nodes. */
int rt_scene_get_nodes(rt_context *ctx, uint32_t scene_id, rt_bvh_node *host_out, uint32_t capacity);
/* Download the leaf-slot -> primitive table (N entries; RT_SPHERE_BIT marks spheres). */
int rt_scene_get_slots(rt_context *ctx, uint32_t scene_id, uint32_t *host_out, uint32_t capacity);
/* Download the wide BVH (nwide nodes of 32*width bytes; root is node 0). */
int rt_scene_get_wide(rt_context *ctx, uint32_t scene_id, void *host_out, size_t capacity_bytes);

/* ---------------- ray queries ---------------- */

#define RT_TRACE_CLOSEST 0u
#define RT_TRACE_ANY 1u        /* any-hit: terminate at the first hit found in [tmin, tmax] */
/* Which primitive a ray hits is decided by the reference's fp32 triangleHit / sphereHit
 * (raytraceBVH.comp:118-181).  By default a hit's t, u, v are then recomputed once in double
 * precision from the primitive's stored fp32 geometry and the fp32 ray (Moller-Trumbore for
 * triangles, the nearest root of the quadratic for spheres) and rounded to fp32, so they lie
 * within an fp32 rounding of the exact intersection.  RT_TRACE_FP32 returns the fp32 values
 * of triangleHit itself instead (bit for bit the reference's arithmetic).  A sphere hit's u, v
 * are the reference's centre-based coordinates (raytraceBVH.comp:172-173) either way. */
#define RT_TRACE_FP32 2u

typedef struct rt_launch_config {
    uint32_t block;        /* threads per block: 64, 128, 256 (0 = tuned, else 128)     */
    uint32_t lds_stack;    /* LDS short-stack entries per lane: 4, 8, 16 (0 = 16)       */
    uint32_t persistent;   /* 1 = persistent work-stealing grid, 0 = one lane per ray    */
    uint32_t refill;       /* persistent: refill a wave when >= this many lanes are idle */
} rt_launch_config;

/* Closest-hit / any-hit ray query (the hitBVH/sceneHit semantics of raytraceBVH.comp:195-274
 * with per-ray [tmin, tmax]).  rays, hits: device.  counts (device, may be NULL): per ray
 * uint32 {AABB tests, primitive tests}.  cfg may be NULL (defaults / autotuned). */
int rt_trace(rt_context *ctx, uint32_t scene_id, const rt_ray *rays, uint32_t n, rt_hit *hits,
             uint32_t *counts, uint32_t flags, const rt_launch_config *cfg, void *stream);

/* ---------------- path tracing (compute stage S2) ---------------- */

#define RT_RENDER_REFERENCE_MATERIALS 1u /* raytraceBVH.comp:100-115: METALLIC/DIELECTRIC absorb */
#define RT_RENDER_NO_PRIMARY_CACHE 2u    /* trace the (identical) primary segment of every sample */
/* How a pixel's samples are spread over lanes.  By default an image with fewer pixels than
 * about four per lane of the device runs every (pixel, sample) on its own lane and sums the
 * samples in order afterwards; larger images run each pixel's samples on one lane.  The
 * image is the same either way. */
#define RT_RENDER_SPLIT_SAMPLES 4u       /* always one lane per (pixel, sample) */
#define RT_RENDER_NO_SPLIT 8u            /* always one lane per pixel */
/* How paths are scheduled.  An image of at least 2^22 pixel-samples is rendered bounce by
 * bounce by default: the primary segments are traced once per pixel, a shading kernel runs
 * every sample's first bounce and queues the paths that scatter, and one persistent launch
 * carries every queued path to its end.  Smaller images carry every path to its end inside
 * one persistent launch from the start.  The image is the same either way.
 * RT_RENDER_SPLIT_SAMPLES and RT_RENDER_NO_SPLIT describe the persistent launch and select it. */
#define RT_RENDER_WAVEFRONT 16u          /* always bounce by bounce */
#define RT_RENDER_PERSISTENT_PATHS 32u   /* always one persistent launch */

/* Clear an RGBA32F accumulation image to (0,0,0,1) (RaytracerBVH.cpp:750-776). */
int rt_accum_clear(rt_context *ctx, float *accum, uint32_t npix, void *stream);

/* spp samples of raytraceBVH.comp's main() per pixel, exactly as spp successive
 * dispatches separated by image barriers (RaytracerBVH.cpp:1025-1050), including the
 * alpha-channel RNG chain.  params is host memory (the UBO).  segments (device, may be
 * NULL): uint64 counter incremented by the number of traced segments. */
int rt_render(rt_context *ctx, uint32_t scene_id, const rt_params *params, uint32_t width,
              uint32_t height, uint32_t spp, uint32_t flags, float *accum, unsigned long long *segments,
              const rt_launch_config *cfg, void *stream);

/* SingleTriangleFullScreen.frag:13-21: out = clamp(sqrt(accum.rgb / spp), 0, 1), a = 1. */
int rt_resolve(rt_context *ctx, const float *accum, uint32_t npix, uint32_t spp, float *out_rgba,
               void *stream);

/* One image of a batched render: the arguments of one rt_render call. */
typedef struct rt_render_job {
    uint32_t scene_id;
    uint32_t width, height, spp;
    uint32_t flags;     /* RT_RENDER_* */
    uint32_t _reserved;
    float *accum;       /* device RGBA32F image, width*height*16 bytes */
    rt_params params;   /* this image's UBO */
} rt_render_job;

/* Renders every job's image: the ones rt_render would run in one persistent launch share
 * one launch per BVH width present; the ones it would render bounce by bounce (see
 * RT_RENDER_WAVEFRONT) follow, each on its own (no such call exists in the reference, whose
 * S2 renders one image: RaytracerBVH.cpp:1000-1078).  Each image is bit-identical to
 * rt_render of its job alone.  jobs: host array; the accum images must not overlap.
 * segments (device, may be NULL) as for rt_render. */
int rt_render_batch(rt_context *ctx, const rt_render_job *jobs, uint32_t njobs, unsigned long long *segments,
                    const rt_launch_config *cfg, void *stream);

/* rt_render without any device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, serving loops that forbid hipMalloc).  The
 * image is bit-identical to rt_render's with the same arguments.  Every scratch buffer (per-sample
 * storage, bounce-by-bounce path state and queues) comes from one static device array of
 * RT_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, instead of the context's cache:
 * bounce-by-bounce images run in as many passes as the array needs, and images that would keep
 * per-sample storage keep whole pixels per lane when it does not fit.  RT_ERR_BUDGET when one
 * sample of a bounce-by-bounce image (120 B per pixel) or a forced RT_RENDER_SPLIT_SAMPLES image
 * (16 B per pixel-sample) exceeds the array.  Calls on one GPU share its array: calls that
 * overlap on one GPU (other streams, other contexts of the process) race. */
#define RT_NOALLOC_SCRATCH_BYTES (256ull << 20)
int rt_render_noalloc(rt_context *ctx, uint32_t scene_id, const rt_params *params, uint32_t width, uint32_t height,
                      uint32_t spp, uint32_t flags, float *accum, unsigned long long *segments,
                      const rt_launch_config *cfg, void *stream);

/* ---------------- batched persistent launch (many scenes, one launch) ---------------- */

typedef struct rt_batch_item {
    uint32_t scene_id;
    uint32_t nrays;
    const rt_ray *rays; /* device */
    rt_hit *hits;       /* device */
} rt_batch_item;

/* All items' rays are traced by one persistent work-stealing launch.  items: host array. */
int rt_trace_batch(rt_context *ctx, const rt_batch_item *items, uint32_t nitems, uint32_t flags,
                   const rt_launch_config *cfg, void *stream);

/* ---------------- autotuner ---------------- */

typedef struct rt_tune_result {
    rt_launch_config cfg;
    float ms;           /* best median time */
    float rays_per_s;
    uint32_t tried;
    uint32_t _pad;
} rt_tune_result;

/* Times launch configurations on (scene, rays) and makes the best the context's default
 * for that scene.  rays/hits: device; the hits buffer is overwritten. */
int rt_autotune(rt_context *ctx, uint32_t scene_id, const rt_ray *rays, uint32_t n, rt_hit *hits,
                uint32_t flags, rt_tune_result *out);

/* Times launch configurations (block size, LDS stack depth, refill threshold) of the fused
 * multi-scene launch on these items and makes the best the context's default for
 * rt_trace_batch calls without a cfg.  The items' hits are overwritten. */
int rt_autotune_batch(rt_context *ctx, const rt_batch_item *items, uint32_t nitems, uint32_t flags,
                      rt_tune_result *out);

/* A sample scene for the build-parameter tuner: host geometry (as for rt_scene_upload,
 * world space) and device rays to trace against it. */
typedef struct rt_tune_scene {
    const rt_triangle *tris;
    uint32_t ntri;
    uint32_t nsph;
    const rt_sphere *spheres;
    const rt_material *materials;
    uint32_t nmat;
    uint32_t nrays;
    const rt_ray *rays; /* device */
    rt_hit *hits;       /* device, overwritten */
} rt_tune_scene;

typedef struct rt_build_candidate {
    uint32_t width, leaf_size;
    float build_ms;         /* device build time summed over the sample scenes */
    float trace_ms;         /* one fused launch over all sample rays, best launch configuration */
    rt_launch_config cfg;   /* that configuration */
    uint64_t device_bytes;  /* resident size of the sample scenes */
} rt_build_candidate;

/* Builds the sample scenes with every BVH width in {2, 4, 8} and leaf size in {1, 2, 4},
 * times a fused launch over all their rays for each (searching launch configurations),
 * and reports every candidate (up to `capacity`, fastest first) and the best build options.
 * The tuner for RaytracerBVH's fixed binary LBVH (RaytracerBVH.cpp:734-997), which has no
 * such parameters.  flags: RT_TRACE_ANY tunes any-hit. */
int rt_autotune_build(rt_context *ctx, const rt_tune_scene *scenes, uint32_t nscenes, uint32_t flags,
                      rt_build_candidate *results, uint32_t capacity, uint32_t *ncandidates,
                      rt_build_options *best);

/* ---------------- serving loop ----------------
 * No counterpart in the reference, which renders one scene per frame on one thread
 * (RaytracerBVH.hpp:517-571).  A server keeps scenes resident (uploaded to its context),
 * queues render and ray-query requests, admits them in arrival order while a slot and
 * enough of its device arena are free, and advances every admitted request in
 * scheduling steps: each step is one fused launch over slices of all admitted requests
 * (samples of renders, rays of queries), so requests that arrive while others run share
 * the device with them.  A request's result does not depend on what it was batched
 * with.  One worker thread per server drives the device; the context's other entry
 * points may be called meanwhile from other threads, but scenes that requests use must
 * not be released until those requests finish (rt_scene_release then reports
 * RT_ERR_BUSY), and servers must be destroyed before their context. */
typedef struct rt_server rt_server;

#define RT_REQ_RENDER 0u
#define RT_REQ_TRACE 1u
#define RT_REQ_OUT_ACCUM 0x100u /* render_flags: return the raw accumulation instead of the resolved image */
#define RT_REQ_DEVICE_BUFFERS 0x200u /* trace_flags: rays and out are device memory; the steps read the rays and
                                       write the hits in place, with no copies (out is required) */

#define RT_SERVER_SERIAL 1u /* one request per step, whole, through the single-scene kernels (the unbatched baseline) */
#define RT_SERVER_NO_YIELD 2u /* device-filling renders run their slices to the end while small requests wait
                                 (by default a slice's last-bounce launch gives the device up to them) */

typedef struct rt_server_config {
    uint64_t arena_bytes;        /* device memory for request buffers: the admission budget (0 = 4 GiB) */
    uint32_t max_active;         /* admitted requests with work not yet enqueued (0 = 256) */
    uint32_t step_us;            /* GPU time a scheduling step may take at most (0 = 50000) */
    uint32_t pipeline;           /* steps enqueued ahead of the device (0 = 2) */
    uint32_t flags;              /* RT_SERVER_* */
    uint32_t max_spp_step;       /* samples per pixel a render advances per step at most (0 = no limit beyond
                                    the step budget); a step holds 16 B per pixel-sample of its renders */
    uint32_t _reserved;
    rt_launch_config render_cfg; /* zeros = defaults */
    rt_launch_config trace_cfg;
} rt_server_config;

typedef struct rt_request {
    uint32_t kind;         /* RT_REQ_RENDER or RT_REQ_TRACE */
    uint32_t scene_id;
    /* render: the arguments of rt_render (accumulation cleared to (0,0,0,1) first) */
    uint32_t width, height, spp;
    uint32_t render_flags; /* RT_RENDER_* | RT_REQ_OUT_ACCUM */
    rt_params params;
    /* ray query */
    uint32_t nrays;
    uint32_t trace_flags;  /* RT_TRACE_ANY | RT_TRACE_FP32 | RT_REQ_DEVICE_BUFFERS */
    const rt_ray *rays;    /* host memory, copied at admission (device memory with RT_REQ_DEVICE_BUFFERS) */
    /* host memory receiving the result when the request completes (NULL = discard):
     * render: width*height RGBA32F; ray query: nrays rt_hit (device memory, written by the
     * query's steps, with RT_REQ_DEVICE_BUFFERS) */
    void *out;
} rt_request;

typedef struct rt_request_stats {
    uint64_t id;
    double submit_ms, admit_ms, done_ms; /* host clock, ms since the server was created */
    uint64_t device_bytes;               /* arena bytes the request held */
    uint32_t steps;                      /* scheduling steps it took part in */
    int32_t status;                      /* RT_OK or the error that failed it */
} rt_request_stats;

typedef struct rt_server_stats {
    uint64_t submitted, completed, failed, steps;
    uint64_t samples;      /* pixel-samples rendered */
    uint64_t segments;     /* path segments traced by renders */
    uint64_t rays;         /* rays traced by ray queries */
    uint64_t arena_bytes, arena_peak;
    double busy_ms;        /* device time of all steps */
    double wall_ms;        /* since the server was created */
    uint32_t active;       /* admitted and not yet complete; may exceed max_active, since a request
                            * whose last slice is enqueued no longer counts against it */
    uint32_t queued;       /* submitted and not yet admitted */
    uint64_t yields;       /* passes of device-filling renders resumed after yielding to small requests */
} rt_server_stats;

int rt_server_create(rt_context *ctx, const rt_server_config *cfg, rt_server **out);
/* Finishes every submitted request, then frees the server. */
int rt_server_destroy(rt_server *srv);
/* Queues a request (RT_ERR_BUDGET if it alone exceeds the arena).  Host arrays are read
 * at admission and `out` is written at completion, so both must stay valid until then. */
int rt_server_submit(rt_server *srv, const rt_request *req, uint64_t *id);
/* Waits for a request (timeout_ms < 0: no limit; RT_ERR_BUSY on timeout).  Returns the
 * request's own status.  stats may be NULL. */
int rt_server_wait(rt_server *srv, uint64_t id, int timeout_ms, rt_request_stats *stats);
/* Waits until every submitted request has finished. */
int rt_server_drain(rt_server *srv);
int rt_server_stats_get(rt_server *srv, rt_server_stats *out);
/* Open-loop replay: submits reqs[i] at arrival_ms[i] (ascending) after the call starts,
 * waits for all of them and writes their stats (n entries; times relative to the call). */
int rt_server_replay(rt_server *srv, const rt_request *reqs, const double *arrival_ms, uint32_t n,
                     rt_request_stats *out);
const char *rt_server_last_error(const rt_server *srv);

/* ---------------- device plumbing for hosts without their own HIP runtime ---------------- */
int rt_util_malloc(void **ptr, size_t bytes);
int rt_util_free(void *ptr);
int rt_util_memcpy_htod(void *dst, const void *src, size_t bytes);
int rt_util_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int rt_util_memcpy_dtod(void *dst, const void *src, size_t bytes);
int rt_util_memset(void *ptr, int value, size_t bytes);
int rt_util_synchronize(void);
int rt_util_stream_create(void **stream);
int rt_util_stream_destroy(void *stream);
int rt_util_stream_synchronize(void *stream);
int rt_util_event_create(void **event);
int rt_util_event_destroy(void *event);
int rt_util_event_record(void *event, void *stream);
int rt_util_event_elapsed_ms(void *start, void *end, float *ms);
int rt_util_device_count(int *count);
int rt_util_device_name(int device, char *buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
