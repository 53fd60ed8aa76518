/*
 * rt_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; never linked by the product).
 *
 * A plain-C restatement of the reference's ray-tracing hot path, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg to check the HIP path's results.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 *
 * Function -> reference source it restates -> what pins it (tests/test_oracle_*.py):
 *   orc_rng_*            random.glsl:10-31           closed-form u32 recurrences (exact golden words),
 *                                                    float-conversion known answers
 *   orc_sinf/orc_cosf    GLSL sin/cos (driver)        libm sin/cos in double over a dense grid (<= 2 ulp),
 *                                                    exact values at 0, symmetries
 *   orc_random_*         random.glsl:33-38,60-62     closed form in double from the same draws
 *   orc_reflect/refract  GLSL builtins (extension)   |r|=|i|, Snell's law, total internal reflection
 *   orc_schlick          textbook (extension)        r(1)=r0, r(0)=1
 *   orc_camera           raytraceBVH.comp:49-81      fp64 pinhole camera (corner/centre rays)
 *   orc_triangle_hit     raytraceBVH.comp:118-149    fp64 Moller-Trumbore (t,u,v 1e-5 rel.), axis-aligned
 *                                                    analytic cases, inclusive edges, parallel cutoff
 *   orc_sphere_hit       raytraceBVH.comp:152-181    analytic sphere distances
 *   orc_sphere_uv        raytraceBVH.comp:172-173    known angles (axis centres), the v = acos(-c.y)/pi
 *                                                    range, NaN outside |c.y| <= 1
 *   orc_refine_*         (not in the reference: the  exact rational arithmetic (Python fractions) on
 *                         ray queries' refined t/u/v)  random and ill-conditioned hits: within one fp32
 *                                                    ulp of the exact value; analytic sphere roots
 *   orc_aabb_hit_ref     raytraceBVH.comp:184-193    analytic slab cases
 *   orc_build_lbvh       GetEnclosingAABB.comp:64-110, GenerateMortonCodesOfPrimitives.comp:41-94,
 *                        RadixSortSimple.comp:52-158, ConstructHLBVH.comp:58-215,
 *                        ConstructAABBsOfInternalNodes.comp:27-65
 *                                                    bit-interleave closed form, stability vs Python's
 *                                                    stable sort, tree invariants (2N-1 nodes, each
 *                                                    reachable once, contiguous leaf ranges, boxes =
 *                                                    unions), hand-built 4-primitive Karras tree
 *   orc_closest (all methods) hitBVH :195-265, raytrace.comp:167-190
 *                                                    brute force == reference traversal == wide replay
 *   orc_render           raytraceBVH.comp:276-375    closed-form single-bounce cases (emitter seen
 *                                                    directly; miss -> 0), RNG chaining across spp
 *   orc_scatter          raytraceBVH.comp:100-115 + the METALLIC/DIELECTRIC extension (A5):
 *                                                    closed-form scenes decided in float64 from the
 *                                                    pinned RNG draws (fuzz-0 mirror into an emitter,
 *                                                    normal incidence through a slab, total internal
 *                                                    reflection), unit-length / Snell / fuzz-cone /
 *                                                    draw-count invariants; planted mistakes in this
 *                                                    function fail them (test_oracle_mutants.py)
 *   orc_resolve          SingleTriangleFullScreen.frag:13-21   closed form
 *   orc_transform_mat4   TransformComponent.cpp:9-45 rotation-matrix identities (orthonormal, det=1)
 *   orc_model_to_world   ModelSpaceToWorldSpace.comp:32-48     identity / translation cases
 */
#include "rt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* fp32 vector helpers: one IEEE operation each, GLSL evaluation order         */
/* ------------------------------------------------------------------------- */
typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
static inline void st3(float *p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline v3 add3(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub3(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul3(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 smul(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 sdiv(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static inline v3 neg3(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot3(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross3(v3 a, v3 b) {
    return mk(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline v3 normalize3(v3 a) { return sdiv(a, sqrtf(dot3(a, a))); }
static inline float gmin(float x, float y) { return (y < x) ? y : x; } /* GLSL min */
static inline float gmax(float x, float y) { return (x < y) ? y : x; } /* GLSL max */
static inline int has_nan3(v3 a) { return a.x != a.x || a.y != a.y || a.z != a.z; }

uint32_t orc_f2u_sat(float x) {
    if (!(x > 0.0f)) return 0u;             /* negative, zero and NaN */
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;                     /* truncation toward zero */
}

/* ------------------------------------------------------------------------- */
/* RNG (random.glsl:10-31)                                                    */
/* ------------------------------------------------------------------------- */
uint32_t orc_rng_seed(uint32_t x, uint32_t y, uint32_t randomState) {
    return (600u * x + y) * (randomState + 1u);
}
uint32_t orc_rng_step(uint32_t s) { return s * 747796405u + 1u; }
float orc_rng_next(uint32_t *state) {
    uint32_t s = orc_rng_step(*state);
    *state = s;
    uint32_t word = ((s >> ((s >> 28) + 4u)) ^ s) * 277803737u;
    word = (word >> 22) ^ word;
    return (float)word / 4294967295.0f; /* the literal rounds to 2^32 in fp32 */
}

/* ------------------------------------------------------------------------- */
/* sin / cos: Cody-Waite reduction by pi/2 (3-part constant) + Cephes sinf/    */
/* cosf minimax kernels on [-pi/4, pi/4].  Valid for |x| < 2^14.              */
/* ------------------------------------------------------------------------- */
static const float kTwoOverPi = 0x1.45f306p-1f;
static const float kPio2A = 0x1.92p+0f;        /* 8 significant bits  */
static const float kPio2B = 0x1.fb6p-12f;      /* 12 significant bits */
static const float kPio2C = -0x1.777a5cp-25f;
static const float kS1 = -0x1.555546p-3f, kS2 = 0x1.11073cp-7f, kS3 = -0x1.9943f2p-13f;
static const float kC1 = 0x1.55554ap-5f, kC2 = -0x1.6c0c34p-10f, kC3 = 0x1.99eb9cp-16f;

static inline float ksin(float r) {
    float z = r * r;
    float p = (kS3 * z + kS2) * z + kS1;
    return r + (r * z) * p;
}
static inline float kcos(float r) {
    float z = r * r;
    float p = (kC3 * z + kC2) * z + kC1;
    return (1.0f - 0.5f * z) + (z * z) * p;
}
static inline float reduce_pio2(float x, int *q) {
    float kf = rintf(x * kTwoOverPi);
    *q = (int)kf;
    return ((x - kf * kPio2A) - kf * kPio2B) - kf * kPio2C;
}
float orc_sinf(float x) {
    int q;
    float r = reduce_pio2(x, &q);
    switch (q & 3) {
    case 0: return ksin(r);
    case 1: return kcos(r);
    case 2: return -ksin(r);
    default: return -kcos(r);
    }
}
float orc_cosf(float x) {
    int q;
    float r = reduce_pio2(x, &q);
    switch (q & 3) {
    case 0: return kcos(r);
    case 1: return -ksin(r);
    case 2: return -kcos(r);
    default: return ksin(r);
    }
}

static const float kPi = 3.1415926535897932385f; /* definitions/random.glsl `pi` */

static v3 rand_in_unit_sphere(uint32_t *st) {
    float rho = orc_rng_next(st);
    float theta = 0.0f + (2.0f * kPi - 0.0f) * orc_rng_next(st); /* random(0, 2*pi) */
    float phi = 0.0f + (kPi - 0.0f) * orc_rng_next(st);          /* random(0, pi)   */
    float sp = orc_sinf(phi);
    return mk(rho * sp * orc_cosf(theta), rho * sp * orc_sinf(theta), rho * orc_cosf(phi));
}
void orc_random_in_unit_sphere(uint32_t *state, float out[3]) { st3(out, rand_in_unit_sphere(state)); }
void orc_random_unit_vector(uint32_t *state, float out[3]) {
    st3(out, normalize3(rand_in_unit_sphere(state)));
}

/* GLSL reflect / refract (textbook; used by the METALLIC / DIELECTRIC extension) */
static inline v3 reflect3(v3 i, v3 n) {
    float dni = dot3(n, i);
    return sub3(i, smul(2.0f * dni, n));
}
static inline v3 refract3(v3 i, v3 n, float eta) {
    float dni = dot3(n, i);
    float k = 1.0f - (eta * eta) * (1.0f - dni * dni);
    if (k < 0.0f) return mk(0.0f, 0.0f, 0.0f);
    return sub3(smul(eta, i), smul(eta * dni + sqrtf(k), n));
}
static inline float schlick(float cosine, float ri) {
    float r0 = (1.0f - ri) / (1.0f + ri);
    r0 = r0 * r0;
    float x = 1.0f - cosine;
    float x5 = ((x * x) * (x * x)) * x;
    return r0 + (1.0f - r0) * x5;
}
void orc_reflect(const float i[3], const float n[3], float out[3]) { st3(out, reflect3(ld3(i), ld3(n))); }
void orc_refract(const float i[3], const float n[3], float eta, float out[3]) {
    st3(out, refract3(ld3(i), ld3(n), eta));
}
float orc_schlick(float cosine, float ri) { return schlick(cosine, ri); }

/* ------------------------------------------------------------------------- */
/* Camera (raytraceBVH.comp:49-81, getRay :329-342)                           */
/* ------------------------------------------------------------------------- */
static const float kFocal = 10.0f; /* _FOCAL_DISTANCE */

void orc_camera(const orc_params *p, uint32_t width, uint32_t height, float out[12]) {
    float W = (float)width, H = (float)height;
    float aspect = W / H;
    float theta = p->verticalFOV * 0.017453292519943295f; /* radians() */
    float h = (float)tan((double)(theta / 2.0f));
    float vh = 2.0f * h * kFocal;
    float vw = vh * aspect;
    v3 pos = ld3(p->camPos), at = ld3(p->camLookAt), up = ld3(p->camUpDir);
    v3 w = normalize3(sub3(pos, at));
    v3 u = normalize3(cross3(up, w));
    v3 v = cross3(w, u);
    v3 vpu = smul(vw, u);
    v3 vpv = smul(vh, neg3(v));
    v3 du = sdiv(vpu, W), dv = sdiv(vpv, H);
    v3 ul = sub3(sub3(sub3(pos, smul(kFocal, w)), sdiv(vpu, 2.0f)), sdiv(vpv, 2.0f));
    v3 p00 = add3(ul, smul(0.5f, add3(du, dv)));
    st3(out + 0, p00);
    st3(out + 3, du);
    st3(out + 6, dv);
    st3(out + 9, pos); /* _DEFOCUS_ANGLE == 0: origin = camPos */
}

void orc_primary_ray(const float cam[12], uint32_t x, uint32_t y, float o[3], float d[3]) {
    v3 p00 = ld3(cam), du = ld3(cam + 3), dv = ld3(cam + 6), org = ld3(cam + 9);
    v3 ps = add3(add3(p00, smul((float)x, du)), smul((float)y, dv)); /* no jitter (A7) */
    st3(o, org);
    st3(d, normalize3(sub3(ps, org)));
}

/* ------------------------------------------------------------------------- */
/* Intersection                                                               */
/* ------------------------------------------------------------------------- */
int orc_triangle_hit(const orc_triangle *tri, const float o_[3], const float d_[3], float tmin,
                     float tmax, float *t_out, float *u_out, float *v_out, float normal[3],
                     int *backface) {
    v3 v0 = ld3(tri->v0), v1 = ld3(tri->v1), v2 = ld3(tri->v2);
    v3 o = ld3(o_), d = ld3(d_);
    v3 u = sub3(v1, v0);
    v3 v = sub3(v2, v0);
    v3 nun = cross3(u, v);
    float nn = dot3(nun, nun);
    if (!(nn > 0.0f)) return 0; /* degenerate triangle: never hit (project convention) */
    v3 n = normalize3(nun);
    float D = dot3(n, v0);
    v3 w = sdiv(nun, nn);
    float denom = dot3(n, d);
    if (fabsf(denom) < 0.0001f) return 0;
    float t = (D - dot3(n, o)) / denom;
    if (t < tmin || t > tmax) return 0;
    v3 p = add3(o, smul(t, d));
    v3 pop = sub3(p, v0);
    float a = dot3(w, cross3(pop, v));
    float b = dot3(w, cross3(u, pop));
    if (a < 0.0f || b < 0.0f || a + b > 1.0f) return 0;
    *t_out = t;
    if (u_out) *u_out = a;
    if (v_out) *v_out = b;
    int bf = dot3(d, n) > 0.0f ? 1 : 0;
    if (backface) *backface = bf;
    if (normal) st3(normal, smul((float)(1 - 2 * bf), n));
    return 1;
}

int orc_sphere_hit(const orc_sphere *s, const float o_[3], const float d_[3], float tmin, float tmax,
                   float *t_out, float normal[3], int *backface) {
    v3 c = ld3(s->center), o = ld3(o_), d = ld3(d_);
    v3 oc = sub3(o, c);
    float a = dot3(d, d);
    float hb = dot3(oc, d);
    float cc = dot3(oc, oc) - (s->radius * s->radius);
    float disc = (hb * hb) - (a * cc);
    if (disc < 0.0f) return 0;
    float sq = sqrtf(disc);
    float root = (-hb - sq) / a;
    if (root < tmin || root > tmax) {
        root = (-hb + sq) / a;
        if (root < tmin || root > tmax) return 0;
    }
    *t_out = root;
    v3 p = add3(o, smul(root, d));
    v3 n = sdiv(sub3(p, c), s->radius);
    int bf = dot3(d, n) > 0.0f ? 1 : 0;
    if (backface) *backface = bf;
    if (normal) st3(normal, smul((float)(1 - 2 * bf), n));
    return 1;
}

/* sphereHit's u, v (raytraceBVH.comp:172-173) depend on the centre only (A17):
 * u = (atan(-c.z, c.x) + pi) / (2 pi), v = acos(-c.y) / pi, evaluated in double, rounded once. */
void orc_sphere_uv(const float c[3], float *u, float *v) {
    const double pi = 3.14159265358979323846;
    *u = (float)((atan2(-(double)c[2], (double)c[0]) + pi) / (2.0 * pi));
    *v = (float)(acos(-(double)c[1]) / pi);
}

/* ------------------------------------------------------------------------- */
/* Refined ray-query results (the product's default, not the reference's):   */
/* t, u, v of a hit recomputed in double from the fp32 primitive and ray.      */
/* Every double operation rounds once (built with -ffp-contract=off).          */
/* ------------------------------------------------------------------------- */
typedef struct { double x, y, z; } d3;
static inline d3 dmk(const float *p) { d3 r; r.x = p[0]; r.y = p[1]; r.z = p[2]; return r; }
static inline d3 dsub(d3 a, d3 b) { d3 r; r.x = a.x - b.x; r.y = a.y - b.y; r.z = a.z - b.z; return r; }
static inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 dcross(d3 a, d3 b) {
    d3 r;
    r.x = a.y * b.z - b.y * a.z;
    r.y = a.z * b.x - b.z * a.x;
    r.z = a.x * b.y - b.x * a.y;
    return r;
}

/* Moller-Trumbore (1997) in double; u = weight of v1, v = weight of v2 (triangleHit's a, b).
 * Returns 0 and leaves the outputs alone when the determinant is exactly zero. */
int orc_refine_triangle(const orc_triangle *tri, const float o_[3], const float d_[3], float *t, float *u,
                        float *v) {
    d3 v0 = dmk(tri->v0), e1 = dsub(dmk(tri->v1), v0), e2 = dsub(dmk(tri->v2), v0);
    d3 o = dmk(o_), d = dmk(d_);
    d3 p = dcross(d, e2);
    double det = ddot(e1, p);
    if (det == 0.0) return 0;
    d3 sv = dsub(o, v0);
    d3 q = dcross(sv, e1);
    *u = (float)(ddot(sv, p) / det);
    *v = (float)(ddot(d, q) / det);
    *t = (float)(ddot(e2, q) / det);
    return 1;
}

/* The root of |o + t d - c|^2 = r^2 nearest the fp32 t, from the stable pair q/a, c/q
 * (Numerical Recipes 5.6).  Returns 0 and leaves t alone for a negative discriminant. */
int orc_refine_sphere(const orc_sphere *s, const float o_[3], const float d_[3], float *t) {
    d3 oc = dsub(dmk(o_), dmk(s->center)), d = dmk(d_);
    double a = ddot(d, d), hb = ddot(oc, d), r = s->radius;
    double cc = ddot(oc, oc) - r * r;
    double disc = hb * hb - a * cc;
    if (!(disc >= 0.0) || !(a > 0.0)) return 0;
    double sq = sqrt(disc);
    double q = hb >= 0.0 ? -(hb + sq) : (sq - hb);
    double r1 = q / a, r2 = q != 0.0 ? cc / q : r1, tf = *t;
    *t = (float)(fabs(r1 - tf) <= fabs(r2 - tf) ? r1 : r2);
    return 1;
}

void orc_refine_hit(const orc_scene *sc, const orc_ray *ray, orc_hit *h) {
    if (h->prim == ORC_NO_HIT) return;
    if (h->prim & ORC_SPHERE_BIT) orc_refine_sphere(sc->sph + (h->prim & ~ORC_SPHERE_BIT), ray->o, ray->d, &h->t);
    else orc_refine_triangle(sc->tris + h->prim, ray->o, ray->d, &h->t, &h->u, &h->v);
}

int orc_aabb_hit_ref(const float o[3], const float d[3], const float bmin[3], const float bmax[3]) {
    float tn = -INFINITY, tf = INFINITY;
    float t1[3], t2[3];
    for (int k = 0; k < 3; ++k) {
        float a = (bmin[k] - o[k]) / d[k];
        float b = (bmax[k] - o[k]) / d[k];
        t1[k] = gmin(a, b);
        t2[k] = gmax(a, b);
    }
    tn = gmax(gmax(t1[0], t1[1]), t1[2]);
    tf = gmin(gmin(t2[0], t2[1]), t2[2]);
    return tn < tf;
}

/* ------------------------------------------------------------------------- */
/* Scene transform (TransformComponent.cpp:9-45, ModelSpaceToWorldSpace.comp) */
/* ------------------------------------------------------------------------- */
void orc_transform_mat4(const float tr[3], const float sc[3], const float rot[3], float m[16]) {
    const float c3 = cosf(rot[2]), s3 = sinf(rot[2]);
    const float c2 = cosf(rot[0]), s2 = sinf(rot[0]);
    const float c1 = cosf(rot[1]), s1 = sinf(rot[1]);
    m[0] = sc[0] * (c1 * c3 + s1 * s2 * s3);
    m[1] = sc[0] * (c2 * s3);
    m[2] = sc[0] * (c1 * s2 * s3 - c3 * s1);
    m[3] = 0.0f;
    m[4] = sc[1] * (c3 * s1 * s2 - c1 * s3);
    m[5] = sc[1] * (c2 * c3);
    m[6] = sc[1] * (c1 * c3 * s2 + s1 * s3);
    m[7] = 0.0f;
    m[8] = sc[2] * (c2 * s1);
    m[9] = sc[2] * (-s2);
    m[10] = sc[2] * (c1 * c2);
    m[11] = 0.0f;
    m[12] = tr[0];
    m[13] = tr[1];
    m[14] = tr[2];
    m[15] = 1.0f;
}

static void mat_vec1(const float *m, const float v[4], float out[4]) {
    float x = v[0], y = v[1], z = v[2];
    for (int i = 0; i < 4; ++i) out[i] = m[0 + i] * x + m[4 + i] * y + m[8 + i] * z + m[12 + i] * 1.0f;
}

void orc_model_to_world(orc_triangle *tris, uint32_t ntri, orc_sphere *sph, uint32_t nsph,
                        const float *models) {
    for (uint32_t i = 0; i < ntri; ++i) {
        const float *m = models + 16u * tris[i].modelIndex;
        float r[4];
        mat_vec1(m, tris[i].v0, r); memcpy(tris[i].v0, r, sizeof r);
        mat_vec1(m, tris[i].v1, r); memcpy(tris[i].v1, r, sizeof r);
        mat_vec1(m, tris[i].v2, r); memcpy(tris[i].v2, r, sizeof r);
    }
    for (uint32_t i = 0; i < nsph; ++i) {
        const float *m = models + 16u * sph[i].modelIndex;
        float r[4];
        mat_vec1(m, sph[i].center, r);
        memcpy(sph[i].center, r, sizeof r); /* radius is not scaled (A17) */
    }
}

/* ------------------------------------------------------------------------- */
/* LBVH build                                                                 */
/* ------------------------------------------------------------------------- */
static inline v3 prim_centroid(const orc_triangle *tris, uint32_t ntri, const orc_sphere *sph, uint32_t i) {
    if (i < ntri) {
        const orc_triangle *t = tris + i;
        v3 s = add3(add3(ld3(t->v0), ld3(t->v1)), ld3(t->v2));
        return sdiv(s, 3.0f);
    }
    return ld3(sph[i - ntri].center);
}
/* total-order min/max (-0 < +0) so the result does not depend on reduction order */
static inline float tmin_total(float a, float b) {
    if (b < a) return b;
    if (a == b && signbit(b) && !signbit(a)) return b;
    return a;
}
static inline float tmax_total(float a, float b) {
    if (a < b) return b;
    if (a == b && !signbit(b) && signbit(a)) return b;
    return a;
}

void orc_enclosing_aabb(const orc_triangle *tris, uint32_t ntri, const orc_sphere *sph, uint32_t nsph,
                        float emin[3], float emax[3]) {
    float mn[3] = {1000000000.0f, 1000000000.0f, 1000000000.0f};
    float mx[3] = {-1000000000.0f, -1000000000.0f, -1000000000.0f};
    uint32_t n = ntri + nsph;
    for (uint32_t i = 0; i < n; ++i) {
        v3 c = prim_centroid(tris, ntri, sph, i);
        float cc[3] = {c.x, c.y, c.z};
        for (int k = 0; k < 3; ++k) {
            mn[k] = tmin_total(mn[k], cc[k]);
            mx[k] = tmax_total(mx[k], cc[k]);
        }
    }
    const float delta = 0.001f, pad = 0.001f / 2.0f;
    for (int k = 0; k < 3; ++k) {
        if (mx[k] - mn[k] < delta) {
            mn[k] -= pad;
            mx[k] += pad;
        }
        emin[k] = mn[k];
        emax[k] = mx[k];
    }
}

static inline uint32_t sep3(uint32_t v) {
    if (v == 1024u) v--;
    v = (v | (v << 16)) & 50331903u;
    v = (v | (v << 8)) & 50393103u;
    v = (v | (v << 4)) & 51130563u;
    v = (v | (v << 2)) & 153391689u;
    return v;
}

void orc_morton_codes(const orc_triangle *tris, uint32_t ntri, const orc_sphere *sph, uint32_t nsph,
                      const float emin[3], const float emax[3], uint32_t flags, uint32_t *codes) {
    uint32_t n = ntri + nsph;
    float span[3] = {emax[0] - emin[0], emax[1] - emin[1], emax[2] - emin[2]};
    for (uint32_t i = 0; i < n; ++i) {
        v3 c = prim_centroid(tris, ntri, sph, i);
        float cc[3] = {c.x, c.y, c.z};
        uint32_t q[3];
        for (int k = 0; k < 3; ++k) {
            float off = (flags & ORC_BUILD_REFERENCE_MORTON) ? cc[k] / span[k] : (cc[k] - emin[k]) / span[k];
            q[k] = orc_f2u_sat(off * 1024.0f);
        }
        codes[i] = (sep3(q[2]) << 2) | (sep3(q[1]) << 1) | sep3(q[0]);
    }
}

void orc_stable_radix_sort(uint32_t *keys, uint32_t *vals, uint32_t n) {
    uint32_t *k2 = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *v2 = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
    for (int pass = 0; pass < 4; ++pass) {
        uint32_t shift = 8u * (uint32_t)pass;
        uint32_t cnt[257];
        memset(cnt, 0, sizeof cnt);
        for (uint32_t i = 0; i < n; ++i) cnt[((ka[i] >> shift) & 255u) + 1u]++;
        for (int b = 0; b < 256; ++b) cnt[b + 1] += cnt[b];
        for (uint32_t i = 0; i < n; ++i) {
            uint32_t b = (ka[i] >> shift) & 255u;
            uint32_t dst = cnt[b]++;
            kb[dst] = ka[i];
            vb[dst] = va[i];
        }
        uint32_t *t;
        t = ka; ka = kb; kb = t;
        t = va; va = vb; vb = t;
    }
    /* 4 passes: the result is back in the caller's arrays */
    free(k2);
    free(v2);
}

static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

/* ConstructHLBVH.comp:64-68 */
static inline int karras_delta(const uint32_t *codes, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    uint32_t ci = codes[i], cj = codes[j];
    if (ci == cj) return 32 + clz32((uint32_t)i ^ (uint32_t)j);
    return clz32(ci ^ cj);
}

static void karras_range(const uint32_t *codes, int n, int id, int *lower, int *upper) {
    int dl = karras_delta(codes, n, id, id - 1);
    int dr = karras_delta(codes, n, id, id + 1);
    int dir = (dr >= dl) ? 1 : -1;
    int dmin = dl < dr ? dl : dr;
    int lmax = 2;
    while (karras_delta(codes, n, id, id + lmax * dir) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t > 0; t >>= 1)
        if (karras_delta(codes, n, id, id + (l + t) * dir) > dmin) l += t;
    int e = id + l * dir;
    *lower = id < e ? id : e;
    *upper = id < e ? e : id;
}

static int karras_split(const uint32_t *codes, int n, int first, int last) {
    int common = karras_delta(codes, n, first, last);
    int split = first;
    int stride = last - first;
    do {
        stride = (stride + 1) >> 1;
        int ns = split + stride;
        if (ns < last) {
            if (karras_delta(codes, n, first, ns) > common) split = ns;
        }
    } while (stride > 1);
    return split;
}

static void prim_box(const orc_triangle *tris, uint32_t ntri, const orc_sphere *sph, uint32_t prim,
                     orc_node *nd) {
    if (prim < ntri) {
        const orc_triangle *t = tris + prim;
        nd->minX = gmin(t->v0[0], gmin(t->v1[0], t->v2[0]));
        nd->maxX = gmax(t->v0[0], gmax(t->v1[0], t->v2[0]));
        nd->minY = gmin(t->v0[1], gmin(t->v1[1], t->v2[1]));
        nd->maxY = gmax(t->v0[1], gmax(t->v1[1], t->v2[1]));
        nd->minZ = gmin(t->v0[2], gmin(t->v1[2], t->v2[2]));
        nd->maxZ = gmax(t->v0[2], gmax(t->v1[2], t->v2[2]));
        nd->primitiveIndex = prim;
        nd->primitiveType = ORC_TRIANGLE_PRIMITIVE;
    } else {
        const orc_sphere *s = sph + (prim - ntri);
        float l[3], r[3];
        for (int k = 0; k < 3; ++k) {
            l[k] = s->center[k] - s->radius;
            r[k] = s->center[k] + s->radius;
        }
        nd->minX = gmin(l[0], r[0]); nd->maxX = gmax(l[0], r[0]);
        nd->minY = gmin(l[1], r[1]); nd->maxY = gmax(l[1], r[1]);
        nd->minZ = gmin(l[2], r[2]); nd->maxZ = gmax(l[2], r[2]);
        nd->primitiveIndex = prim - ntri;
        nd->primitiveType = ORC_SPHERE_PRIMITIVE;
    }
    const float delta = 0.001f, pad = 0.001f / 2.0f;
    if (nd->maxX - nd->minX < delta) { nd->minX -= pad; nd->maxX += pad; }
    if (nd->maxY - nd->minY < delta) { nd->minY -= pad; nd->maxY += pad; }
    if (nd->maxZ - nd->minZ < delta) { nd->minZ -= pad; nd->maxZ += pad; }
    nd->left = 0;
    nd->right = 0;
}

int orc_build_lbvh(const orc_triangle *tris, uint32_t ntri, const orc_sphere *sph, uint32_t nsph,
                   uint32_t flags, orc_node *nodes, uint32_t *parent, uint32_t *slot_prim,
                   uint32_t *sorted_codes, uint32_t *sorted_prims) {
    uint32_t n = ntri + nsph;
    if (n == 0) return 0;
    float emin[3], emax[3];
    orc_enclosing_aabb(tris, ntri, sph, nsph, emin, emax);
    uint32_t *codes = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *prims = (uint32_t *)malloc(sizeof(uint32_t) * n);
    orc_morton_codes(tris, ntri, sph, nsph, emin, emax, flags, codes);
    for (uint32_t i = 0; i < n; ++i) prims[i] = i;
    orc_stable_radix_sort(codes, prims, n);
    const int N = (int)n, leaf0 = N - 1;

    /* leaves (ConstructHLBVH.comp:152-169; A11 selects unsorted placement) */
    for (int i = 0; i < N; ++i) {
        uint32_t p = (flags & ORC_BUILD_REFERENCE_LEAVES) ? (uint32_t)i : prims[i];
        prim_box(tris, ntri, sph, p, &nodes[leaf0 + i]);
        slot_prim[i] = p < ntri ? p : (ORC_SPHERE_BIT | (p - ntri));
    }
    if (parent) parent[0] = 0;
    /* internal nodes (ConstructHLBVH.comp:171-214) */
    for (int i = 0; i < N - 1; ++i) {
        int first, last;
        karras_range(codes, N, i, &first, &last);
        int split = karras_split(codes, N, first, last);
        int left = (split == first) ? leaf0 + split : split;
        int right = (split + 1 == last) ? leaf0 + split + 1 : split + 1;
        orc_node *nd = &nodes[i];
        nd->minX = nd->maxX = nd->minY = nd->maxY = nd->minZ = nd->maxZ = 0.0f;
        nd->left = (uint32_t)left;
        nd->right = (uint32_t)right;
        nd->primitiveIndex = 0;
        nd->primitiveType = 0;
        if (parent) {
            parent[left] = (uint32_t)i;
            parent[right] = (uint32_t)i;
        }
    }
    /* refit: post-order (ConstructAABBsOfInternalNodes.comp:27-65 computes the same unions) */
    if (N > 1) {
        uint32_t *stack = (uint32_t *)malloc(sizeof(uint32_t) * 2u * n);
        uint8_t *done = (uint8_t *)calloc(n, 1);
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            uint32_t id = stack[sp - 1];
            orc_node *nd = &nodes[id];
            int lpend = nd->left < (uint32_t)leaf0 && !done[nd->left];
            int rpend = nd->right < (uint32_t)leaf0 && !done[nd->right];
            if (lpend || rpend) {
                if (rpend) stack[sp++] = nd->right;
                if (lpend) stack[sp++] = nd->left;
                continue;
            }
            const orc_node *a = &nodes[nd->left], *b = &nodes[nd->right];
            nd->minX = gmin(a->minX, b->minX); nd->maxX = gmax(a->maxX, b->maxX);
            nd->minY = gmin(a->minY, b->minY); nd->maxY = gmax(a->maxY, b->maxY);
            nd->minZ = gmin(a->minZ, b->minZ); nd->maxZ = gmax(a->maxZ, b->maxZ);
            done[id] = 1;
            sp--;
        }
        free(stack);
        free(done);
    }
    if (sorted_codes) memcpy(sorted_codes, codes, sizeof(uint32_t) * n);
    if (sorted_prims) memcpy(sorted_prims, prims, sizeof(uint32_t) * n);
    free(codes);
    free(prims);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Closest / any hit                                                          */
/* ------------------------------------------------------------------------- */
typedef struct { float t, u, v; v3 normal; uint32_t material; uint32_t prim; int backface; } hitrec;

static int test_prim(const orc_scene *sc, uint32_t prim, v3 o, v3 d, float tmin, float tmax, hitrec *h) {
    float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    float t, u = 0.0f, v = 0.0f, n[3];
    int bf;
    if (prim & ORC_SPHERE_BIT) {
        const orc_sphere *s = sc->sph + (prim & ~ORC_SPHERE_BIT);
        if (!orc_sphere_hit(s, oo, dd, tmin, tmax, &t, n, &bf)) return 0;
        orc_sphere_uv(s->center, &u, &v);
        h->material = s->materialIndex;
    } else {
        const orc_triangle *tr = sc->tris + prim;
        if (!orc_triangle_hit(tr, oo, dd, tmin, tmax, &t, &u, &v, n, &bf)) return 0;
        h->material = tr->materialIndex;
    }
    h->t = t;
    h->u = u;
    h->v = v;
    h->normal = ld3(n);
    h->prim = prim;
    h->backface = bf;
    return 1;
}

/* robustness factor for the wide traversal's slab test: 1 + 2*gamma(3) rounded up */
static const float kRobust = 0x1.000006p+0f;

#define WIDE_EMPTY 0xFFFFFFFFu
#define WIDE_LEAF 0x80000000u

typedef struct { uint32_t e; float tn; } wentry;

/* order-preserving float -> uint32 map (total order, -0 < +0) */
static inline uint32_t ord_key(float f) {
    uint32_t b;
    memcpy(&b, &f, sizeof b);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

/* The distance a pushed entry carries: the kernel's stack keeps only the upper 16 bits of
 * ord_key (traverse.hpp push/pop), which decode to a distance no larger than tn. */
static inline float stacked_tn(float tn) {
    uint32_t k = ord_key(tn) & 0xFFFF0000u, b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    float f;
    memcpy(&f, &b, sizeof f);
    return f;
}

static uint32_t closest_wide(const orc_scene *sc, v3 o, v3 d, float tmin, float tmax, hitrec *best,
                             orc_counts *cnt, int any) {
    const uint32_t W = sc->width;
    const uint32_t stride = 8u * W;
    const uint32_t *nodes = (const uint32_t *)sc->wide_nodes;
    /* slab distances fma(plane, idir, n_lo or n_hi): n = -(o * idir) rounded once per ray, lowered for
     * the planes an axis enters through and raised for those it leaves through by 4u|n|; a zero
     * direction component takes idir = +-2^100 and no offset (traverse.hpp slab_axis) */
    float idir[3], nlo[3], nhi[3];
    const float dd[3] = {d.x, d.y, d.z}, oo[3] = {o.x, o.y, o.z};
    for (int k = 0; k < 3; ++k) {
        idir[k] = dd[k] != 0.0f ? 1.0f / dd[k] : copysignf(0x1p100f, dd[k]);
        const float n = -(oo[k] * idir[k]);
        const float e = dd[k] != 0.0f ? copysignf(0x1p-22f * fabsf(n), idir[k]) : 0.0f;
        nlo[k] = n - e;
        nhi[k] = n + e;
    }
    float bt = tmax;
    uint32_t bslot = ORC_NO_HIT;
    wentry stack[1024];
    int sp = 0;
    if (sc->nwide == 0) return ORC_NO_HIT;
    /* The kernel's per-lane state machine (traverse.hpp): a current entry and the stack.  A node step
     * or a leaf step ends by popping when nothing is current. */
    wentry cur = {sc->wide_root, tmin};
    int have = 1;
#define ORC_POP()                                                                 \
    do {                                                                          \
        if (!have && sp > 0) { cur = stack[--sp]; have = 1; }                     \
    } while (0)
    while (have) {
        if (cur.e & WIDE_LEAF) { /* leaf step */
            have = 0;
            if (cur.tn <= bt * kRobust) {
                uint32_t first = cur.e & 0x07FFFFFFu, count = ((cur.e >> 27) & 15u) + 1u;
                for (uint32_t s = first; s < first + count; ++s) {
                    hitrec h;
                    if (cnt) cnt->n_tri++;
                    if (test_prim(sc, sc->slot_prim[s], o, d, tmin, bt, &h)) {
                        if (h.t < bt || (h.t == bt && s < bslot)) {
                            bt = h.t;
                            bslot = s;
                            *best = h;
                            if (any) return bslot;
                        }
                    }
                }
            }
            ORC_POP();
            continue;
        }
        if (cur.tn > bt * kRobust) { /* node culled by a closer hit found since it was pushed */
            have = 0;
            ORC_POP();
            continue;
        }
        const uint32_t *nd = nodes + (size_t)stride * cur.e;
        const float *f = (const float *)nd;
        wentry cand[8];
        int nh = 0;
        for (uint32_t c = 0; c < W; ++c) {
            uint32_t ch = nd[6u * W + c];
            if (ch == WIDE_EMPTY) continue;
            if (cnt) cnt->n_box++;
            float x0 = fmaf(f[0 * W + c], idir[0], nlo[0]), x1 = fmaf(f[1 * W + c], idir[0], nhi[0]);
            float y0 = fmaf(f[2 * W + c], idir[1], nlo[1]), y1 = fmaf(f[3 * W + c], idir[1], nhi[1]);
            float z0 = fmaf(f[4 * W + c], idir[2], nlo[2]), z1 = fmaf(f[5 * W + c], idir[2], nhi[2]);
            float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
            float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), bt));
            if (tn <= tf * kRobust) {
                cand[nh].e = ch;
                cand[nh].tn = tn;
                nh++;
            }
        }
        if (!any) { /* ascending (order-preserving key of the entry distance, entry) */
            for (int i = 1; i < nh; ++i) {
                wentry x = cand[i];
                uint64_t kx = ((uint64_t)ord_key(x.tn) << 32) | x.e;
                int j = i - 1;
                while (j >= 0 && (((uint64_t)ord_key(cand[j].tn) << 32) | cand[j].e) > kx) {
                    cand[j + 1] = cand[j];
                    j--;
                }
                cand[j + 1] = x;
            }
        } else {
            for (int i = 0; i < nh; ++i) cand[i].tn = -INFINITY; /* any-hit entries are never culled */
        }
        /* the nearest becomes the current entry at once, with its full distance; the others go
         * through the kernel's stack, farthest first */
        for (int i = nh - 1; i >= 1; --i) {
            wentry e = cand[i];
            if (!any) e.tn = stacked_tn(e.tn);
            stack[sp++] = e;
        }
        if (nh > 0) cur = cand[0];
        else have = 0;
        if (sp > 1000) abort();
        ORC_POP();
    }
#undef ORC_POP
    return bslot;
}

static uint32_t closest_ref_bvh(const orc_scene *sc, v3 o, v3 d, float tmin, float tmax, hitrec *best,
                                orc_counts *cnt) {
    const uint32_t N = sc->nslots;
    if (N == 0) return ORC_NO_HIT;
    float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    float closest = tmax;
    uint32_t bslot = ORC_NO_HIT;
    uint32_t stack[4096];
    uint32_t sp = 0, cur = 0;
    for (;;) {
        const orc_node *nd = &sc->nodes[cur];
        float bmin[3] = {nd->minX, nd->minY, nd->minZ}, bmax[3] = {nd->maxX, nd->maxY, nd->maxZ};
        if (cnt) cnt->n_box++;
        if (orc_aabb_hit_ref(oo, dd, bmin, bmax)) {
            if (nd->left == 0 && nd->right == 0) {
                uint32_t prim = nd->primitiveType == ORC_SPHERE_PRIMITIVE ? (ORC_SPHERE_BIT | nd->primitiveIndex)
                                                                          : nd->primitiveIndex;
                hitrec h;
                if (cnt) cnt->n_tri++;
                if (test_prim(sc, prim, o, d, tmin, closest, &h)) {
                    closest = h.t;
                    *best = h;
                    bslot = cur - (N - 1u);
                }
                if (sp == 0) break;
                cur = stack[--sp];
            } else {
                if (sp >= 4096) abort();
                stack[sp++] = nd->left;
                cur = nd->right;
            }
        } else {
            if (sp == 0) break;
            cur = stack[--sp];
        }
    }
    return bslot;
}

static uint32_t closest_brute(const orc_scene *sc, v3 o, v3 d, float tmin, float tmax, hitrec *best,
                              orc_counts *cnt) {
    float closest = tmax;
    uint32_t bslot = ORC_NO_HIT;
    /* descending slot order with replace-on-equal == the reference's right-first visit order */
    for (uint32_t k = sc->nslots; k-- > 0;) {
        hitrec h;
        if (cnt) cnt->n_tri++;
        if (test_prim(sc, sc->slot_prim[k], o, d, tmin, closest, &h)) {
            closest = h.t;
            bslot = k;
            *best = h;
        }
    }
    return bslot;
}

static uint32_t closest_any_method(const orc_scene *sc, int method, v3 o, v3 d, float tmin, float tmax,
                                   hitrec *h, orc_counts *cnt) {
    if (has_nan3(d)) return ORC_NO_HIT; /* NaN direction: the root slab test fails */
    switch (method) {
    case ORC_METHOD_REF_BVH: return closest_ref_bvh(sc, o, d, tmin, tmax, h, cnt);
    case ORC_METHOD_WIDE: return closest_wide(sc, o, d, tmin, tmax, h, cnt, 0);
    default: return closest_brute(sc, o, d, tmin, tmax, h, cnt);
    }
}

uint32_t orc_closest(const orc_scene *sc, int method, const float o[3], const float d[3], float tmin,
                     float tmax, orc_hit *hit, float normal[3], uint32_t *material, orc_counts *counts) {
    hitrec h;
    uint32_t s = closest_any_method(sc, method, ld3(o), ld3(d), tmin, tmax, &h, counts);
    if (s == ORC_NO_HIT) {
        if (hit) { hit->prim = ORC_NO_HIT; hit->t = tmax; hit->u = 0.0f; hit->v = 0.0f; }
        return s;
    }
    if (hit) { hit->prim = h.prim; hit->t = h.t; hit->u = h.u; hit->v = h.v; }
    if (normal) st3(normal, h.normal);
    if (material) *material = h.material;
    return s;
}

int orc_any(const orc_scene *sc, int method, const float o_[3], const float d_[3], float tmin, float tmax,
            orc_hit *hit, orc_counts *counts) {
    v3 o = ld3(o_), d = ld3(d_);
    hitrec h;
    int found = 0;
    if (!has_nan3(d)) {
        if (method == ORC_METHOD_WIDE) {
            found = closest_wide(sc, o, d, tmin, tmax, &h, counts, 1) != ORC_NO_HIT;
        } else {
            for (uint32_t k = 0; k < sc->nslots && !found; ++k) {
                if (counts) counts->n_tri++;
                found = test_prim(sc, sc->slot_prim[k], o, d, tmin, tmax, &h);
            }
        }
    }
    if (hit) {
        if (found) { hit->prim = h.prim; hit->t = h.t; hit->u = h.u; hit->v = h.v; }
        else { hit->prim = ORC_NO_HIT; hit->t = tmax; hit->u = 0.0f; hit->v = 0.0f; }
    }
    return found;
}

void orc_trace_batch(const orc_scene *sc, int method, const orc_ray *rays, uint32_t n, uint32_t flags,
                     orc_hit *hits, orc_counts *counts, int threads) {
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_num_procs();
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
#endif
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const orc_ray *r = rays + i;
        orc_counts c = {0, 0};
        if (flags & 1u) orc_any(sc, method, r->o, r->d, r->tmin, r->tmax, &hits[i], &c);
        else orc_closest(sc, method, r->o, r->d, r->tmin, r->tmax, &hits[i], NULL, NULL, &c);
        if (!(flags & ORC_TRACE_FP32)) orc_refine_hit(sc, r, &hits[i]);
        if (counts) counts[i] = c;
    }
    (void)threads;
}

/* ------------------------------------------------------------------------- */
/* Path tracer                                                                */
/* ------------------------------------------------------------------------- */
static v3 ray_color(const orc_scene *sc, int method, v3 o, v3 dir, uint32_t depth, uint32_t flags,
                    uint32_t *st, uint64_t *segs) {
    v3 color = mk(0.0f, 0.0f, 0.0f);
    v3 atten = mk(1.0f, 1.0f, 1.0f);
    v3 d = normalize3(dir);
    const v3 background = mk(0.0f, 0.0f, 0.0f);
    for (uint32_t i = 0; i < depth; ++i) {
        hitrec h;
        if (segs) (*segs)++;
        uint32_t s = closest_any_method(sc, method, o, d, 0.001f, 10000000.0f, &h, NULL);
        if (s == ORC_NO_HIT) {
            color = add3(color, mul3(background, atten));
            break;
        }
        const orc_material *m = sc->mats + h.material;
        v3 alb = ld3(m->albedo);
        v3 p = add3(o, smul(h.t, d));
        v3 emitted = m->materialType == ORC_LIGHT ? alb : mk(0.0f, 0.0f, 0.0f);
        color = add3(color, mul3(emitted, atten));
        float dd[3], nn[3], att[3], nd[3];
        st3(dd, d);
        st3(nn, h.normal);
        if (!orc_scatter(m, flags, dd, nn, h.backface, st, att, nd)) break;
        atten = mul3(atten, ld3(att));
        o = p;
        d = ld3(nd);
    }
    return color;
}

/* scatter() of raytraceBVH.comp:100-115 (DIFFUSE; LIGHT and, with ORC_RENDER_REFERENCE_MATERIALS,
 * METALLIC/DIELECTRIC absorb) and the textbook METALLIC/DIELECTRIC of Ray Tracing in One Weekend
 * (secs. 10-11), the extension A5: fuzzed mirror reflection, absorbed when it points into the surface;
 * Snell refraction with Schlick's reflectance, total internal reflection drawing no random number. */
int orc_scatter(const orc_material *m, uint32_t flags, const float d_[3], const float n_[3], int backface,
                uint32_t *st, float att_[3], float out[3]) {
    v3 d = ld3(d_), n = ld3(n_), alb = ld3(m->albedo);
    v3 att = mk(0.0f, 0.0f, 0.0f), nd = mk(0.0f, 0.0f, 0.0f);
    int scattered = 0;
    if (m->materialType == ORC_DIFFUSE) {
        att = alb;
        v3 ruv = normalize3(rand_in_unit_sphere(st));
        nd = normalize3(add3(n, ruv));
        scattered = 1;
    } else if (m->materialType == ORC_METALLIC && !(flags & ORC_RENDER_REFERENCE_MATERIALS)) {
        att = alb;
        v3 refl = reflect3(d, n);
        v3 rs = rand_in_unit_sphere(st);
        nd = normalize3(add3(refl, smul(m->albedo[3], rs)));
        scattered = dot3(nd, n) > 0.0f;
    } else if (m->materialType == ORC_DIELECTRIC && !(flags & ORC_RENDER_REFERENCE_MATERIALS)) {
        att = mk(1.0f, 1.0f, 1.0f);
        float ior = m->albedo[3];
        float ri = backface ? ior : (1.0f / ior); /* front face: entering the medium */
        float ct = gmin(dot3(neg3(d), n), 1.0f);
        float stt = sqrtf(1.0f - ct * ct);
        int cannot = ri * stt > 1.0f;
        if (cannot || schlick(ct, ri) > orc_rng_next(st)) nd = reflect3(d, n);
        else nd = refract3(d, n, ri);
        nd = normalize3(nd);
        scattered = 1;
    }
    st3(att_, att);
    st3(out, nd);
    return scattered;
}

void orc_render_rows(const orc_scene *sc, int method, const orc_params *p, uint32_t width, uint32_t height,
                     uint32_t spp, uint32_t flags, float *accum, uint32_t row0, uint32_t row1,
                     uint64_t *segments, int threads) {
    float cam[12];
    orc_camera(p, width, height, cam);
    uint64_t total = 0;
    if (row1 > height) row1 = height;
    int64_t npix = (int64_t)(row1 - row0) * width;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_num_procs();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads) reduction(+ : total)
#endif
    for (int64_t k = 0; k < npix; ++k) {
        uint32_t y = row0 + (uint32_t)(k / width), x = (uint32_t)(k % width);
        float *acc = accum + 4u * ((size_t)y * width + x);
        const uint32_t seed0 = orc_rng_seed(x, y, p->randomState);
        uint64_t segs = 0;
        for (uint32_t s = 0; s < spp; ++s) {
            uint32_t st = seed0 + orc_f2u_sat(acc[3] * 4294967294.0f); /* A2/A3 */
            float next = orc_rng_next(&st);
            float o[3], d[3];
            orc_primary_ray(cam, x, y, o, d);
            v3 c = ray_color(sc, method, ld3(o), ld3(d), p->maxRayTraceDepth, flags, &st, &segs);
            acc[0] = c.x + acc[0];
            acc[1] = c.y + acc[1];
            acc[2] = c.z + acc[2];
            acc[3] = next;
        }
        total += segs;
    }
    if (segments) *segments = total;
    (void)threads;
}

void orc_render(const orc_scene *sc, int method, const orc_params *p, uint32_t width, uint32_t height,
                uint32_t spp, uint32_t flags, float *accum, uint64_t *segments, int threads) {
    orc_render_rows(sc, method, p, width, height, spp, flags, accum, 0, height, segments, threads);
}

void orc_resolve(const float *accum, uint32_t npix, uint32_t spp, float *out) {
    for (uint32_t i = 0; i < npix; ++i) {
        for (int k = 0; k < 3; ++k) {
            float x = sqrtf(accum[4u * i + k] / (float)spp);
            out[4u * i + k] = gmin(gmax(x, 0.0f), 1.0f);
        }
        out[4u * i + 3] = 1.0f;
    }
}
