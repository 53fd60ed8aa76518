// rt_device.hpp -- device-side arithmetic of the ray-tracing hot path (gfx950).
//
// Every fp32 operation below is one IEEE round-to-nearest operation in the
// evaluation order of the reference GLSL (RaytracerGPU_MastersProject/shaders/),
// never contracted to FMA: the library is compiled with -ffp-contract=off, and
// HIP's division and sqrt are correctly rounded by default.  This fixes the
// arithmetic the reference leaves to the Vulkan driver so results are
// reproducible bit for bit on any IEEE machine.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace rtgpu {

struct v3 {
    float x, y, z;
};

__host__ __device__ inline v3 mk3(float x, float y, float z) { return v3{x, y, z}; }
__host__ __device__ inline v3 operator+(v3 a, v3 b) { return v3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline v3 operator-(v3 a, v3 b) { return v3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline v3 operator-(v3 a) { return v3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline v3 mul(v3 a, v3 b) { return v3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ inline v3 operator*(float s, v3 a) { return v3{s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline v3 operator/(v3 a, float s) { return v3{a.x / s, a.y / s, a.z / s}; }
// GLSL dot: (a.x*b.x + a.y*b.y) + a.z*b.z
__host__ __device__ inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// GLSL cross (GLSL 4.50 spec 8.5)
__host__ __device__ inline v3 cross(v3 a, v3 b) {
    return v3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
// GLSL normalize(v) = v / length(v)
__host__ __device__ inline v3 normalize(v3 a) { return a / sqrtf(dot(a, a)); }
// GLSL min/max: min(x,y) = y < x ? y : x; max(x,y) = x < y ? y : x
__host__ __device__ inline float gmin(float x, float y) { return (y < x) ? y : x; }
__host__ __device__ inline float gmax(float x, float y) { return (x < y) ? y : x; }
__host__ __device__ inline bool has_nan(v3 a) { return a.x != a.x || a.y != a.y || a.z != a.z; }

// float -> uint conversion, saturating, NaN -> 0 (the AMD v_cvt_u32_f32 rule)
__host__ __device__ inline uint32_t f2u_sat(float x) {
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

// ---------------------------------------------------------------------------
// PCG-style RNG of random.glsl:10-31 (pcg32i step with inc 1, RXS-M-XS output)
// ---------------------------------------------------------------------------
__host__ __device__ inline uint32_t rng_seed(uint32_t x, uint32_t y, uint32_t random_state) {
    return (600u * x + y) * (random_state + 1u);  // the literal 600 aliases rows >= 600
}
__host__ __device__ inline float rng_next(uint32_t &state) {
    uint32_t s = state * 747796405u + 1u;
    state = s;
    uint32_t word = ((s >> ((s >> 28) + 4u)) ^ s) * 277803737u;
    word = (word >> 22) ^ word;
    return (float)word / 4294967295.0f;  // the constant is 2^32 in fp32
}

// ---------------------------------------------------------------------------
// sin / cos: reduction by pi/2 with a three-part Cody-Waite constant (first
// two parts short enough that k*part is exact for the small k used here),
// then the Cephes single-precision minimax kernels on [-pi/4, pi/4].
// ---------------------------------------------------------------------------
namespace trig {
constexpr float two_over_pi = 0x1.45f306p-1f;
constexpr float pio2_a = 0x1.92p+0f;
constexpr float pio2_b = 0x1.fb6p-12f;
constexpr float pio2_c = -0x1.777a5cp-25f;
constexpr float s1 = -0x1.555546p-3f, s2 = 0x1.11073cp-7f, s3 = -0x1.9943f2p-13f;
constexpr float c1 = 0x1.55554ap-5f, c2 = -0x1.6c0c34p-10f, c3 = 0x1.99eb9cp-16f;

__host__ __device__ inline float ksin(float r) {
    float z = r * r;
    float p = (s3 * z + s2) * z + s1;
    return r + (r * z) * p;
}
__host__ __device__ inline float kcos(float r) {
    float z = r * r;
    float p = (c3 * z + c2) * z + c1;
    return (1.0f - 0.5f * z) + (z * z) * p;
}
__host__ __device__ inline float reduce(float x, int &q) {
    float k = rintf(x * two_over_pi);
    q = (int)k;
    return ((x - k * pio2_a) - k * pio2_b) - k * pio2_c;
}
}  // namespace trig

__host__ __device__ inline void sincos_rt(float x, float &s, float &c) {
    int q;
    float r = trig::reduce(x, q);
    float sr = trig::ksin(r), cr = trig::kcos(r);
    switch (q & 3) {
        case 0: s = sr; c = cr; break;
        case 1: s = cr; c = -sr; break;
        case 2: s = -sr; c = -cr; break;
        default: s = -cr; c = sr; break;
    }
}

constexpr float kPi = 3.1415926535897932385f;

// random.glsl:33-38 -- rho, theta, phi each uniform (not volume-uniform)
__host__ __device__ inline v3 random_in_unit_sphere(uint32_t &st) {
    float rho = rng_next(st);
    float theta = 0.0f + (2.0f * kPi - 0.0f) * rng_next(st);
    float phi = 0.0f + (kPi - 0.0f) * rng_next(st);
    float sp, cp, stt, ct;
    sincos_rt(phi, sp, cp);
    sincos_rt(theta, stt, ct);
    return v3{rho * sp * ct, rho * sp * stt, rho * cp};
}

// GLSL reflect/refract builtins (GLSL 4.50 spec 8.5)
__host__ __device__ inline v3 reflect(v3 i, v3 n) {
    float dni = dot(n, i);
    return i - (2.0f * dni) * n;
}
__host__ __device__ inline v3 refract(v3 i, v3 n, float eta) {
    float dni = dot(n, i);
    float k = 1.0f - (eta * eta) * (1.0f - dni * dni);
    if (k < 0.0f) return v3{0.0f, 0.0f, 0.0f};
    return eta * i - (eta * dni + sqrtf(k)) * n;
}
// Schlick's approximation (Ray Tracing in One Weekend, sec. 11.4)
__host__ __device__ inline float schlick(float cosine, float ri) {
    float r0 = (1.0f - ri) / (1.0f + ri);
    r0 = r0 * r0;
    float x = 1.0f - cosine;
    float x5 = ((x * x) * (x * x)) * x;
    return r0 + (1.0f - r0) * x5;
}

// ---------------------------------------------------------------------------
// Leaf records (64 B per leaf slot, slot order = BVH leaf order).
//   triangle: r0 = {v0, D}, r1 = {e1, n.x}, r2 = {e2, n.y}, r3 = {w, n.z}
//     with e1 = v1-v0, e2 = v2-v0, n = normalize(cross(e1,e2)), D = dot(n,v0),
//     w = cross(e1,e2)/dot(cross,cross): the ray-independent part of
//     triangleHit (raytraceBVH.comp:119-125), precomputed once.
//   sphere:   r0 = {center, radius}, r1 = {kSphereTag, u, v, 0} with u, v the
//     centre-based coordinates of sphereHit (raytraceBVH.comp:172-173), one pair per sphere
// ---------------------------------------------------------------------------
constexpr uint32_t kSphereTag = 0xFFFFFFFFu;

struct __attribute__((aligned(16))) LeafRecord {
    float4 r0, r1, r2, r3;
};

__host__ __device__ inline LeafRecord make_triangle_record(v3 v0, v3 v1, v3 v2) {
    v3 e1 = v1 - v0;
    v3 e2 = v2 - v0;
    v3 nun = cross(e1, e2);
    float nn = dot(nun, nun);
    LeafRecord r;
    if (!(nn > 0.0f)) {  // degenerate: n = 0 makes |denom| < 1e-4, i.e. never hit
        r.r0 = make_float4(v0.x, v0.y, v0.z, 0.0f);
        r.r1 = make_float4(e1.x, e1.y, e1.z, 0.0f);
        r.r2 = make_float4(e2.x, e2.y, e2.z, 0.0f);
        r.r3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return r;
    }
    v3 n = normalize(nun);
    float D = dot(n, v0);
    v3 w = nun / nn;
    r.r0 = make_float4(v0.x, v0.y, v0.z, D);
    r.r1 = make_float4(e1.x, e1.y, e1.z, n.x);
    r.r2 = make_float4(e2.x, e2.y, e2.z, n.y);
    r.r3 = make_float4(w.x, w.y, w.z, n.z);
    return r;
}

// Per-ray part of triangleHit (raytraceBVH.comp:126-136).  Returns t (or a
// negative sentinel path via the bool).  t in [tmin, tmax] inclusive.
__device__ inline bool triangle_test(const LeafRecord &r, v3 o, v3 d, float tmin, float tmax, float &t_out,
                                     float &a_out, float &b_out) {
    v3 n = v3{r.r1.w, r.r2.w, r.r3.w};
    float denom = dot(n, d);
    if (fabsf(denom) < 0.0001f) return false;
    float t = (r.r0.w - dot(n, o)) / denom;
    if (t < tmin || t > tmax) return false;
    v3 p = o + t * d;
    v3 pop = p - v3{r.r0.x, r.r0.y, r.r0.z};
    v3 e1 = v3{r.r1.x, r.r1.y, r.r1.z};
    v3 e2 = v3{r.r2.x, r.r2.y, r.r2.z};
    v3 w = v3{r.r3.x, r.r3.y, r.r3.z};
    float a = dot(w, cross(pop, e2));
    float b = dot(w, cross(e1, pop));
    if (a < 0.0f || b < 0.0f || a + b > 1.0f) return false;
    t_out = t;
    a_out = a;
    b_out = b;
    return true;
}

// sphereHit (raytraceBVH.comp:152-170): nearest root in [tmin, tmax], else far root
__device__ inline bool sphere_test(const LeafRecord &r, v3 o, v3 d, float tmin, float tmax, float &t_out) {
    v3 c = v3{r.r0.x, r.r0.y, r.r0.z};
    float radius = r.r0.w;
    v3 oc = o - c;
    float a = dot(d, d);
    float hb = dot(oc, d);
    float cc = dot(oc, oc) - (radius * radius);
    float disc = (hb * hb) - (a * cc);
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float root = (-hb - sq) / a;
    if (root < tmin || root > tmax) {
        root = (-hb + sq) / a;
        if (root < tmin || root > tmax) return false;
    }
    t_out = root;
    return true;
}

__device__ inline bool is_sphere_record(const LeafRecord &r) { return __float_as_uint(r.r1.x) == kSphereTag; }

// sphereHit's u, v (raytraceBVH.comp:172-173) are computed from the sphere's centre, not the
// hit point, so they are one pair per sphere: u = (atan(-c.z, c.x) + pi) / (2 pi),
// v = acos(-c.y) / pi, here evaluated in double and rounded once (NaN when |c.y| > 1, as acos).
inline void sphere_uv(const float c[3], float &u, float &v) {
    const double pi = 3.14159265358979323846;
    u = (float)((atan2(-(double)c[2], (double)c[0]) + pi) / (2.0 * pi));
    v = (float)(acos(-(double)c[1]) / pi);
}

// ---------------------------------------------------------------------------
// Refined ray-query results.  triangleHit decides which primitive a ray hits, in fp32, as
// the reference does; its t, u, v carry the rounding of that plane-based form (on
// 0.005-unit triangles seen from a unit away, ~1e-5 relative).  A ray query then
// recomputes them once per hit ray in double precision from the stored fp32 vertices and
// the fp32 ray, and rounds once: within half an fp32 ulp of the exact intersection
// except where double rounding or the conditioning of the triangle intervene.
//   triangle: Moller-Trumbore (Moller & Trumbore 1997), u = weight of v1, v = weight of v2,
//             the convention of triangleHit's a and b;
//   sphere:   the root of the half-b quadratic nearest the fp32 t, with the numerically
//             stable pair q/a, c/q (Press et al., Numerical Recipes 5.6).
// Every double operation rounds once (no contraction), so the CPU restatement reproduces
// the results bit for bit.
// ---------------------------------------------------------------------------
struct d3 {
    double x, y, z;
};
__host__ __device__ inline d3 dsub(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline d3 dcross(d3 a, d3 b) {
    return d3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y};
}
__host__ __device__ inline d3 dvec(float x, float y, float z) { return d3{(double)x, (double)y, (double)z}; }

// false (outputs untouched) when the determinant is exactly zero
__host__ __device__ inline bool refine_triangle(d3 v0, d3 v1, d3 v2, d3 o, d3 d, float &t, float &u, float &v) {
    const d3 e1 = dsub(v1, v0), e2 = dsub(v2, v0);
    const d3 p = dcross(d, e2);
    const double det = ddot(e1, p);
    if (det == 0.0) return false;
    const d3 s = dsub(o, v0);
    const d3 q = dcross(s, e1);
    u = (float)(ddot(s, p) / det);
    v = (float)(ddot(d, q) / det);
    t = (float)(ddot(e2, q) / det);
    return true;
}

// false (t untouched) when the discriminant is negative in double precision
__host__ __device__ inline bool refine_sphere(d3 c, double radius, d3 o, d3 d, float &t) {
    const d3 oc = dsub(o, c);
    const double a = ddot(d, d);
    const double hb = ddot(oc, d);
    const double cc = ddot(oc, oc) - radius * radius;
    const double disc = hb * hb - a * cc;
    if (!(disc >= 0.0) || !(a > 0.0)) return false;
    const double sq = sqrt(disc);
    const double q = hb >= 0.0 ? -(hb + sq) : (sq - hb);
    const double r1 = q / a;
    const double r2 = q != 0.0 ? cc / q : r1;
    const double tf = (double)t;
    const double pick = fabs(r1 - tf) <= fabs(r2 - tf) ? r1 : r2;
    t = (float)pick;
    return true;
}

// ---------------------------------------------------------------------------
// Wide BVH node: W children, SoA inside the node so a lane fetches its node with
// 16-byte vector loads.  Layout in 32-bit words (stride 8*W words = 32*W bytes):
//   [0W) lo.x  [1W) hi.x  [2W) lo.y  [3W) hi.y  [4W) lo.z  [5W) hi.z  [6W) child  [7W) pad
// child: kEmpty, a wide node index (bit 31 clear), or a leaf
//   kLeafBit | (count-1) << 27 | first_slot.
// ---------------------------------------------------------------------------
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kSlotMask = 0x07FFFFFFu;
// slab-test robustness: exit distance scaled by 1 + 2*gamma(3) (Ize 2013), rounded up
constexpr float kRobust = 0x1.000006p+0f;

__host__ __device__ inline uint32_t leaf_entry(uint32_t first, uint32_t count) {
    return kLeafBit | ((count - 1u) << 27) | first;
}
__host__ __device__ inline uint32_t leaf_first(uint32_t e) { return e & kSlotMask; }
__host__ __device__ inline uint32_t leaf_count(uint32_t e) { return ((e >> 27) & 15u) + 1u; }

}  // namespace rtgpu
