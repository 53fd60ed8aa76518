/*
 * rt_oracle.h -- CPU oracle for the BVH ray-tracing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * GLSL hot path (RaytracerGPU_MastersProject/shaders/compute/raytraceBVH.comp,
 * shaders/include/random.glsl, the LBVH build shaders and the fragment resolve).
 * It exists to check the HIP product path; nothing in the product links it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Arithmetic convention (the reference leaves it to the Vulkan driver):
 *   - every fp32 operation is a single IEEE-754 round-to-nearest operation,
 *     evaluated left to right as written in the GLSL, never contracted to FMA
 *     (build with -ffp-contract=off);
 *   - dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z; cross per the GLSL spec;
 *   - normalize(v) = v / sqrt(dot(v,v)) with IEEE sqrt and division;
 *   - GLSL min(x,y) = (y < x) ? y : x, max(x,y) = (x < y) ? y : x;
 *   - float -> uint conversion saturates (NaN -> 0), as AMD hardware does;
 *   - sin/cos are the project's own fp32 functions (Cody-Waite reduction by
 *     pi/2 in three parts, Cephes single-precision minimax kernels).
 *
 * Parity pins: see the header of rt_oracle.c, which lists per function what
 * pins it (closed forms, invariants, brute force, libm in double precision).
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- byte layouts of the reference's SSBOs / UBO (definitions.glsl, SceneTypes.hpp) ---- */

typedef struct { /* definitions.glsl:15-21, SceneTypes.hpp:41-47: 64 B */
    float v0[4];
    float v1[4];
    float v2[4];
    uint32_t materialIndex;
    uint32_t modelIndex;
    uint32_t _pad[2];
} orc_triangle;

typedef struct { /* definitions.glsl:23-28, SceneTypes.hpp:62-66: 32 B */
    float center[4];
    float radius;
    uint32_t materialIndex;
    uint32_t modelIndex;
    uint32_t _pad;
} orc_sphere;

typedef struct { /* definitions.glsl:10-13, SceneTypes.hpp:82-84: 32 B; albedo[3] = fuzz / ior extension */
    float albedo[4];
    uint32_t materialType;
    uint32_t _pad[3];
} orc_material;

typedef struct { /* definitions.glsl:52-72, SceneTypes.hpp:107-123: 40 B */
    float minX, maxX, minY, maxY, minZ, maxZ;
    uint32_t left, right, primitiveIndex, primitiveType;
} orc_node;

typedef struct { /* raytraceBVH.comp:7-18 (std140), RaytracerBVH.hpp:31-42: 80 B */
    float camPos[4];
    float camLookAt[4];
    float camUpDir[4];
    float verticalFOV;
    uint32_t numTriangles;
    uint32_t numSpheres;
    uint32_t numMaterials;
    uint32_t numLights;
    uint32_t maxRayTraceDepth;
    uint32_t randomState;
    uint32_t _pad;
} orc_params;

typedef struct { float o[3]; float tmin; float d[3]; float tmax; } orc_ray;      /* 32 B */
typedef struct { uint32_t prim; float t; float u; float v; } orc_hit;            /* 16 B */

enum { ORC_LIGHT = 0, ORC_DIFFUSE = 1, ORC_METALLIC = 2, ORC_DIELECTRIC = 3 };
enum { ORC_SPHERE_PRIMITIVE = 0, ORC_TRIANGLE_PRIMITIVE = 1 };
#define ORC_NO_HIT 0xFFFFFFFFu
#define ORC_SPHERE_BIT 0x80000000u

/* ---- scalar building blocks (exported so tests can pin them one by one) ---- */
uint32_t orc_rng_seed(uint32_t x, uint32_t y, uint32_t randomState);          /* random.glsl:10 */
uint32_t orc_rng_step(uint32_t s);                                            /* random.glsl:13-15 */
float    orc_rng_next(uint32_t *state);                                       /* random.glsl:18-23 */
float    orc_sinf(float x);
float    orc_cosf(float x);
uint32_t orc_f2u_sat(float x);
void     orc_random_in_unit_sphere(uint32_t *state, float out[3]);             /* random.glsl:33-38 */
void     orc_random_unit_vector(uint32_t *state, float out[3]);                /* random.glsl:60-62 */
void     orc_reflect(const float i[3], const float n[3], float out[3]);
void     orc_refract(const float i[3], const float n[3], float eta, float out[3]);
float    orc_schlick(float cosine, float ri);

/* camera (raytraceBVH.comp:49-81): out = pixel00[3], deltaU[3], deltaV[3], origin[3] */
void orc_camera(const orc_params *p, uint32_t width, uint32_t height, float out[12]);
void orc_primary_ray(const float cam[12], uint32_t x, uint32_t y, float o[3], float d[3]);

/* intersection (raytraceBVH.comp:118-193). Return 1 on hit and fill t/u/v/normal(face-forwarded)/backface */
int orc_triangle_hit(const orc_triangle *tri, const float o[3], const float d[3], float tmin, float tmax,
                     float *t, float *u, float *v, float normal[3], int *backface);
int orc_sphere_hit(const orc_sphere *s, const float o[3], const float d[3], float tmin, float tmax,
                   float *t, float normal[3], int *backface);
int orc_aabb_hit_ref(const float o[3], const float d[3], const float bmin[3], const float bmax[3]);
/* raytraceBVH.comp:172-173: a sphere hit's u, v from its centre */
void orc_sphere_uv(const float center[3], float *u, float *v);
/* refined results (the product's default ray-query output; no reference counterpart):
 * t/u/v recomputed in double and rounded once; 0 = no solution in double, outputs untouched */
int orc_refine_triangle(const orc_triangle *tri, const float o[3], const float d[3], float *t, float *u,
                        float *v);
int orc_refine_sphere(const orc_sphere *s, const float o[3], const float d[3], float *t);

/* ---- LBVH builder (GetEnclosingAABB, GenerateMortonCodesOfPrimitives, RadixSortSimple,
 *      ConstructHLBVH, ConstructAABBsOfInternalNodes) ---- */
#define ORC_BUILD_REFERENCE_LEAVES 1u  /* A11: leaf N-1+i holds primitive i (unsorted)   */
#define ORC_BUILD_REFERENCE_MORTON 2u  /* A12: quantise coord/span instead of (coord-min)/span */

void orc_model_to_world(orc_triangle *tris, uint32_t ntri, orc_sphere *sph, uint32_t nsph,
                        const float *models /* n x 16, column-major mat4 */);
void orc_transform_mat4(const float translation[3], const float scale[3], const float rotation[3],
                        float out[16]);
void orc_enclosing_aabb(const orc_triangle *tris, uint32_t ntri, const orc_sphere *sph, uint32_t nsph,
                        float emin[3], float emax[3]);
void orc_morton_codes(const orc_triangle *tris, uint32_t ntri, const orc_sphere *sph, uint32_t nsph,
                      const float emin[3], const float emax[3], uint32_t flags, uint32_t *codes);
void orc_stable_radix_sort(uint32_t *keys, uint32_t *vals, uint32_t n);
/* Full build. nodes: 2N-1 entries; parent: 2N-1 entries (construction info .parent);
 * slot_prim: N entries, leaf slot -> primitive id (triangle i, or ORC_SPHERE_BIT|j).
 * sorted_codes / sorted_prims (optional, may be NULL): the sort output. Returns 0 on success. */
int orc_build_lbvh(const orc_triangle *tris, uint32_t ntri, const orc_sphere *sph, uint32_t nsph,
                   uint32_t flags, orc_node *nodes, uint32_t *parent, uint32_t *slot_prim,
                   uint32_t *sorted_codes, uint32_t *sorted_prims);

/* ---- closest / any hit over one ray ---- */
typedef struct {
    const orc_triangle *tris; uint32_t ntri;
    const orc_sphere *sph; uint32_t nsph;
    const orc_material *mats; uint32_t nmat;
    const orc_node *nodes; uint32_t nnodes;        /* binary LBVH (reference builder output) */
    const uint32_t *slot_prim; uint32_t nslots;    /* leaf slot -> primitive */
    /* wide BVH as exported by the product (for traversal replay), may be NULL */
    const void *wide_nodes; uint32_t nwide; uint32_t width; uint32_t wide_root;
} orc_scene;

enum { ORC_METHOD_BRUTE = 0, ORC_METHOD_REF_BVH = 1, ORC_METHOD_WIDE = 2 };

typedef struct { uint32_t n_box; uint32_t n_tri; } orc_counts;

/* closest hit; returns slot (or ORC_NO_HIT), fills hit (prim id, t, u, v) and shading data */
uint32_t orc_closest(const orc_scene *sc, int method, const float o[3], const float d[3],
                     float tmin, float tmax, orc_hit *hit, float normal[3], uint32_t *material,
                     orc_counts *counts);
int orc_any(const orc_scene *sc, int method, const float o[3], const float d[3], float tmin,
            float tmax, orc_hit *hit, orc_counts *counts);

/* batch ray queries; flags: bit0 = any-hit, ORC_TRACE_FP32 = the fp32 triangleHit t/u/v (default:
 * refined as orc_refine_*). counts may be NULL. threads <= 0 -> all cores */
#define ORC_TRACE_FP32 2u
void orc_trace_batch(const orc_scene *sc, int method, const orc_ray *rays, uint32_t n, uint32_t flags,
                     orc_hit *hits, orc_counts *counts, int threads);

/* ---- path tracer (raytraceBVH.comp:276-375) ---- */
#define ORC_RENDER_REFERENCE_MATERIALS 1u /* A5: METALLIC/DIELECTRIC absorb (reference behaviour) */
/* accum: W*H*4 floats, in/out (RGBA32F; caller clears to (0,0,0,1) like RaytracerBVH.cpp:772).
 * Runs spp samples per pixel exactly as spp successive dispatches would.
 * segments (optional): number of closest-hit queries issued. */
void orc_render(const orc_scene *sc, int method, const orc_params *p, uint32_t width, uint32_t height,
                uint32_t spp, uint32_t flags, float *accum, uint64_t *segments, int threads);
void orc_render_rows(const orc_scene *sc, int method, const orc_params *p, uint32_t width,
                     uint32_t height, uint32_t spp, uint32_t flags, float *accum, uint32_t row0,
                     uint32_t row1, uint64_t *segments, int threads);
/* SingleTriangleFullScreen.frag:13-21 */
void orc_resolve(const float *accum, uint32_t npix, uint32_t spp, float *rgba_out);
/* one bounce's scatter (raytraceBVH.comp:100-115 + the METALLIC/DIELECTRIC extension): d the unit
 * incoming direction, n the face-forwarded unit normal; returns 1 when the path continues along out
 * with attenuation att.  Draws from *state as the path tracer does. */
int orc_scatter(const orc_material *m, uint32_t flags, const float d[3], const float n[3], int backface,
                uint32_t *state, float att[3], float out[3]);

#ifdef __cplusplus
}
#endif
#endif
