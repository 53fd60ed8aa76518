// traverse.hpp -- wide-BVH traversal shared by the ray-query and path-tracing kernels.
//
// One lane owns one ray.  A ray's current entry is either a wide node or a leaf:
//   node_step(): W slab tests; hit children are ordered by (entry distance, entry)
//                -- one 64-bit key each, sorted by a network -- the nearest becomes
//                the current entry and the rest are pushed far-to-near;
//   leaf_step(): the leaf's primitives are tested (triangleHit / sphereHit).
// Each ray's sequence of visited entries depends only on the ray and the tree, so
// its counts of AABB and primitive tests are reproducible.  The kernels choose, per
// wave and per iteration, whether lanes sitting on nodes or lanes sitting on leaves
// advance (a ballot); lanes of the other kind simply wait, which changes when a ray
// advances but never the order of its own steps.  This keeps the ~80-instruction
// triangle test from being executed by the whole wave on almost every iteration.
//
// The stack keeps its newest S entries in LDS columns private to the lane and spills
// older ones to a private array, so deep trees never fault.  A slot is 6 bytes: the
// 32-bit entry and the upper half of its distance's order key (see push()).
//
// Closest-hit semantics are those of hitBVH (raytraceBVH.comp:195-265) for any
// traversal order: the smallest t wins and, among equal t, the smallest leaf slot
// wins -- exactly the primitive the reference's right-first traversal tests last,
// because it accepts equal t (`t > tMax` rejects, :130).
#pragma once

#include "rt_device.hpp"

namespace rtgpu {

constexpr uint32_t kSpillCap = 184;  // private spill entries per lane (beyond the LDS part)
constexpr uint32_t kChunk = 512;     // work items a wave takes from the global counter per atomic

// Which rays a wave takes per atomic.  The counter counts chunks, not rays: the first rays go in
// chunks of `chunk`, the last 10 per resident lane in chunks of 128 and then 64.  When the
// counter runs out, a wave is then left with at most one or two rays per lane instead of a
// whole chunk (8 rays per lane at 512), and the launch drains sooner; the few thousand extra
// atomics cost less than that drain.
// T is the ray-index type: uint32_t for one batch of rays, uint64_t for the fused many-scene
// launch.
template <typename T> struct ChunkSchedule {
    T n, nA, nB, KA, KB;
    uint32_t chunk;

    // shrink: fewer rays per chunk when there are too few for every wave to take two chunks
    // (a small launch then ends with its longest ray, not with a wave's run of rays)
    __device__ __forceinline__ ChunkSchedule(T n_, uint32_t waves, bool shrink) {
        n = n_;
        chunk = shrink ? (uint32_t)max((T)64u, min((T)kChunk, (T)(n / (2u * waves)) & ~(T)63u)) : kChunk;
        const T lanes = (T)waves * 64u;
        const T nC = chunk > 64u ? min(n, (T)(2u * lanes)) : (T)0u;
        nB = chunk > 128u ? min((T)(n - nC), (T)(8u * lanes)) : (T)0u;
        nA = n - nB - nC;
        KA = (nA + chunk - 1u) / chunk;
        KB = (nB + 127u) / 128u;
    }
    // rays [lo, hi) of chunk c; false when c is past the end
    __device__ __forceinline__ bool range(T c, uint64_t &lo, uint64_t &hi) const {
        if (c < KA) {
            lo = (uint64_t)c * chunk;
            hi = min(lo + chunk, (uint64_t)nA);
        } else if (c - KA < KB) {
            lo = (uint64_t)nA + (uint64_t)(c - KA) * 128u;
            hi = min(lo + 128u, (uint64_t)nA + nB);
        } else {
            lo = (uint64_t)nA + nB + (uint64_t)(c - KA - KB) * 64u;
            hi = min(lo + 64u, (uint64_t)n);
        }
        return lo < (uint64_t)n;
    }
};

template <int W> struct WideNode {
    float lo[3][W], hi[3][W];
    uint32_t child[W];
};

template <int W> __device__ __forceinline__ void load_node(const uint32_t *wide, uint32_t idx, WideNode<W> &nd);

template <> __device__ __forceinline__ void load_node<2>(const uint32_t *wide, uint32_t idx, WideNode<2> &nd) {
    const float4 *p = reinterpret_cast<const float4 *>(wide) + (size_t)idx * 4u;
    float4 x = p[0], y = p[1], z = p[2];
    uint4 c = reinterpret_cast<const uint4 *>(p)[3];
    nd.lo[0][0] = x.x; nd.lo[0][1] = x.y; nd.hi[0][0] = x.z; nd.hi[0][1] = x.w;
    nd.lo[1][0] = y.x; nd.lo[1][1] = y.y; nd.hi[1][0] = y.z; nd.hi[1][1] = y.w;
    nd.lo[2][0] = z.x; nd.lo[2][1] = z.y; nd.hi[2][0] = z.z; nd.hi[2][1] = z.w;
    nd.child[0] = c.x; nd.child[1] = c.y;
}

__device__ __forceinline__ void unpack4(float4 v, float *o) { o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w; }

template <> __device__ __forceinline__ void load_node<4>(const uint32_t *wide, uint32_t idx, WideNode<4> &nd) {
    const float4 *p = reinterpret_cast<const float4 *>(wide) + (size_t)idx * 8u;
    unpack4(p[0], nd.lo[0]); unpack4(p[1], nd.hi[0]);
    unpack4(p[2], nd.lo[1]); unpack4(p[3], nd.hi[1]);
    unpack4(p[4], nd.lo[2]); unpack4(p[5], nd.hi[2]);
    uint4 c = reinterpret_cast<const uint4 *>(p)[6];
    nd.child[0] = c.x; nd.child[1] = c.y; nd.child[2] = c.z; nd.child[3] = c.w;
}

template <> __device__ __forceinline__ void load_node<8>(const uint32_t *wide, uint32_t idx, WideNode<8> &nd) {
    const float4 *p = reinterpret_cast<const float4 *>(wide) + (size_t)idx * 16u;
    unpack4(p[0], nd.lo[0]); unpack4(p[1], nd.lo[0] + 4);
    unpack4(p[2], nd.hi[0]); unpack4(p[3], nd.hi[0] + 4);
    unpack4(p[4], nd.lo[1]); unpack4(p[5], nd.lo[1] + 4);
    unpack4(p[6], nd.hi[1]); unpack4(p[7], nd.hi[1] + 4);
    unpack4(p[8], nd.lo[2]); unpack4(p[9], nd.lo[2] + 4);
    unpack4(p[10], nd.hi[2]); unpack4(p[11], nd.hi[2] + 4);
    uint4 c0 = reinterpret_cast<const uint4 *>(p)[12], c1 = reinterpret_cast<const uint4 *>(p)[13];
    nd.child[0] = c0.x; nd.child[1] = c0.y; nd.child[2] = c0.z; nd.child[3] = c0.w;
    nd.child[4] = c1.x; nd.child[5] = c1.y; nd.child[6] = c1.z; nd.child[7] = c1.w;
}

// order-preserving map float -> uint32 (total order, -0 < +0) and back
__device__ __forceinline__ uint32_t ord_key(float f) {
    const uint32_t b = __float_as_uint(f);
    return b ^ ((uint32_t)((int32_t)b >> 31) | 0x80000000u);  // negative: ~b; otherwise b | sign bit
}
__device__ __forceinline__ float ord_dec(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

constexpr uint32_t kOrdLowest = 0x00800000u;  // ord_key(-FLT_MAX): never culled, and exact in 16 bits

// per-lane traversal state
struct Trav {
    v3 o, d, idir;
    v3 nlo, nhi;        // per-axis slab offsets of the low and high planes (see slab_axis)
    float tmin, bt;     // ray start; best t so far (initially tmax)
    uint32_t bslot;     // leaf slot of the best hit (kEmpty = none)
    float bu, bv;       // barycentrics of the best hit
    uint32_t cur;       // current entry (node index or leaf entry)
    float cur_tn;       // its entry distance
    bool have;          // cur is valid
    uint32_t top, nspill;
    uint32_t nbox, ntri;
};

// A stack slot holds an entry and the upper 16 bits of ord_key(entry distance).  The
// truncated key decodes to a distance no larger than the true one, so culling a popped
// entry against the best hit stays conservative; it may only keep an entry that the full
// key would have culled.  Six bytes per slot instead of eight let more waves share a
// CU's LDS.
constexpr uint32_t kStackSlotBytes = 6;

struct StackMem {
    uint32_t *ent;    // this lane's LDS entry column: slot k at ent[k * stride]
    uint16_t *key;    // this lane's LDS key column, same indexing
    uint32_t stride;  // block size
    uint2 *spill;     // private spill array (kSpillCap entries of (entry, key << 16))
};

// LDS layout of a block: S x block entries, then S x block keys
template <int S> __device__ __forceinline__ StackMem stack_mem(void *lds, uint2 *spill) {
    uint32_t *ent = static_cast<uint32_t *>(lds);
    uint16_t *key = reinterpret_cast<uint16_t *>(ent + (size_t)S * blockDim.x);
    return StackMem{ent + threadIdx.x, key + threadIdx.x, blockDim.x, spill};
}

inline size_t stack_lds_bytes(uint32_t lds_stack, uint32_t block) {
    return (size_t)kStackSlotBytes * lds_stack * block;
}

template <int S> __device__ __forceinline__ void push(Trav &t, const StackMem &m, uint32_t e, uint32_t tnkey) {
    if (t.top - t.nspill == (uint32_t)S) {  // LDS part full: move its oldest entry out
        const uint32_t o = (t.nspill & (S - 1)) * m.stride;
        if (t.nspill < kSpillCap) m.spill[t.nspill] = make_uint2(m.ent[o], (uint32_t)m.key[o] << 16);
        t.nspill++;
    }
    const uint32_t o = (t.top & (S - 1)) * m.stride;
    m.ent[o] = e;
    m.key[o] = (uint16_t)(tnkey >> 16);
    t.top++;
}

template <int S> __device__ __forceinline__ void pop(Trav &t, const StackMem &m) {
    uint2 v;
    if (t.top > t.nspill) {
        t.top--;
        const uint32_t o = (t.top & (S - 1)) * m.stride;
        v = make_uint2(m.ent[o], (uint32_t)m.key[o] << 16);
    } else {
        t.nspill--;
        t.top--;
        v = (t.nspill < kSpillCap) ? m.spill[t.nspill] : make_uint2(kEmpty, 0u);
    }
    t.cur = v.x;
    t.cur_tn = ord_dec(v.y);
    t.have = v.x != kEmpty;
}

// Slab distances take one FMA per plane: fma(plane, idir, n) with n = -(o * idir) rounded once
// per ray.  That rounding is an absolute error of up to u|n| (u = 2^-24) in every distance of the
// axis, which the relative widening kRobust does not cover, so each axis carries two offsets:
// nlo for the planes it enters through and nhi for those it leaves through, n lowered and raised
// by e = 4u|n| (which also covers the rounding of n -+ e).  Near distances come out no larger and
// far distances no smaller than with the exact offset, and a box the exact test enters is never
// culled.  The bound widens an axis only where that axis decides the entry or exit distance: a
// per-ray widening by max e over the axes would be huge for a ray almost parallel to an axis
// plane (|d_i| of 1e-7 gives |n| of 1e7) and would send it through most of the tree.
// A zero direction component gets idir = +-2^100 instead of infinity, with e = 0: the products
// with a power of two are exact, so that axis's distances are (plane - o) * 2^100 rounded once
// (coordinates below 2^27 in magnitude), the slab of a direction component of 2^-100 with the
// zero's sign.  A ray lying exactly in the plane of a box face on such an axis enters the box where
// that nudged ray moves into it (low face with +0, high face with -0) and is culled on the other
// face; the reference's AABBhitCheck (raytraceBVH.comp:184-193) culls it on both through 0 * inf =
// NaN, and brute force finds what a cull misses.
__device__ __forceinline__ void slab_axis(float d, float o, float &idir, float &nlo, float &nhi) {
    idir = d != 0.0f ? 1.0f / d : copysignf(0x1p100f, d);
    const float n = -(o * idir);
    const float e = d != 0.0f ? copysignf(0x1p-22f * fabsf(n), idir) : 0.0f;
    nlo = n - e;
    nhi = n + e;
}

__device__ __forceinline__ void trav_init(Trav &t, v3 o, v3 d, float tmin, float tmax, uint32_t nwide) {
    t.o = o;
    t.d = d;
    slab_axis(d.x, o.x, t.idir.x, t.nlo.x, t.nhi.x);
    slab_axis(d.y, o.y, t.idir.y, t.nlo.y, t.nhi.y);
    slab_axis(d.z, o.z, t.idir.z, t.nlo.z, t.nhi.z);
    t.tmin = tmin;
    t.bt = tmax;
    t.bslot = kEmpty;
    t.bu = 0.0f;
    t.bv = 0.0f;
    t.cur = 0u;  // root
    t.cur_tn = tmin;
    t.have = nwide > 0 && !has_nan(d);  // a NaN direction misses (the reference's root slab test fails)
    t.top = 0;
    t.nspill = 0;
    t.nbox = 0;
    t.ntri = 0;
}

__device__ __forceinline__ bool trav_done(const Trav &t) { return !t.have && t.top == 0; }
__device__ __forceinline__ bool at_leaf(const Trav &t) { return t.have && (t.cur & kLeafBit); }
__device__ __forceinline__ bool at_node(const Trav &t) { return t.have && !(t.cur & kLeafBit); }

__device__ __forceinline__ void cswap64(uint64_t *k, int i, int j) {
    const uint64_t a = k[i], b = k[j];
    const bool sw = a > b;
    k[i] = sw ? b : a;
    k[j] = sw ? a : b;
}

template <int W> __device__ __forceinline__ void sort_keys(uint64_t *k) {
    if constexpr (W == 2) {
        cswap64(k, 0, 1);
    } else if constexpr (W == 4) {
        cswap64(k, 0, 1); cswap64(k, 2, 3);
        cswap64(k, 0, 2); cswap64(k, 1, 3);
        cswap64(k, 1, 2);
    } else {  // Batcher odd-even merge sort, 19 comparators
        cswap64(k, 0, 1); cswap64(k, 2, 3); cswap64(k, 4, 5); cswap64(k, 6, 7);
        cswap64(k, 0, 2); cswap64(k, 1, 3); cswap64(k, 4, 6); cswap64(k, 5, 7);
        cswap64(k, 1, 2); cswap64(k, 5, 6);
        cswap64(k, 0, 4); cswap64(k, 1, 5); cswap64(k, 2, 6); cswap64(k, 3, 7);
        cswap64(k, 2, 4); cswap64(k, 3, 5);
        cswap64(k, 1, 2); cswap64(k, 3, 4); cswap64(k, 5, 6);
    }
}

// Process the current wide node.  Precondition: at_node(t).
//   closest: hit children in increasing (ord_key(tn), entry); ANY: in child order.
//   COUNT: keep t.nbox / t.ntri (only launches that report per-ray counts need them).
template <int W, int S, bool ANY, bool COUNT = true>
__device__ __forceinline__ void node_step(Trav &t, const StackMem &m, const uint32_t *__restrict__ wide) {
    if (t.cur_tn > t.bt * kRobust) {
        t.have = false;  // culled by a closer hit found since it was pushed
    } else {
        WideNode<W> nd;
        load_node<W>(wide, t.cur, nd);
        uint64_t key[W];
        int nh = 0;
#pragma unroll
        for (int c = 0; c < W; ++c) {
            const bool valid = nd.child[c] != kEmpty;
            if constexpr (COUNT) t.nbox += valid ? 1u : 0u;
            float x0 = fmaf(nd.lo[0][c], t.idir.x, t.nlo.x), x1 = fmaf(nd.hi[0][c], t.idir.x, t.nhi.x);
            float y0 = fmaf(nd.lo[1][c], t.idir.y, t.nlo.y), y1 = fmaf(nd.hi[1][c], t.idir.y, t.nhi.y);
            float z0 = fmaf(nd.lo[2][c], t.idir.z, t.nlo.z), z1 = fmaf(nd.hi[2][c], t.idir.z, t.nhi.z);
            float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), t.tmin));
            float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), t.bt));
            const bool hit = valid && tn <= tf * kRobust;
            nh += hit ? 1 : 0;
            const uint32_t hi = ANY ? (uint32_t)c : ord_key(tn);
            // a missed child's key sorts after every hit child's (ord_key of a distance is below ~0u)
            key[c] = ((uint64_t)(hit ? hi : ~0u) << 32) | nd.child[c];
        }
        sort_keys<W>(key);
        if (nh == 0) {
            t.have = false;
        } else {
#pragma unroll
            for (int k = W - 1; k >= 1; --k)
                if (k < nh) push<S>(t, m, (uint32_t)key[k], ANY ? kOrdLowest : (uint32_t)(key[k] >> 32));
            t.cur = (uint32_t)key[0];
            t.cur_tn = ANY ? -__uint_as_float(0x7F800000u) : ord_dec((uint32_t)(key[0] >> 32));
        }
    }
    if (!t.have && t.top > 0) pop<S>(t, m);
}

// Test the primitives of one leaf entry; true when an any-hit query has found its hit.
template <bool ANY, bool COUNT>
__device__ __forceinline__ bool test_leaf(Trav &t, uint32_t entry, const LeafRecord *__restrict__ rec) {
    const uint32_t first = leaf_first(entry), cnt = leaf_count(entry);
    for (uint32_t s = first; s < first + cnt; ++s) {
        if constexpr (COUNT) t.ntri++;
        const LeafRecord r = rec[s];
        float th = 0.0f, a = 0.0f, b = 0.0f;
        bool hit;
        if (is_sphere_record(r)) {
            hit = sphere_test(r, t.o, t.d, t.tmin, t.bt, th);
            a = r.r1.y;  // the sphere's centre-based u, v
            b = r.r1.z;
        } else {
            hit = triangle_test(r, t.o, t.d, t.tmin, t.bt, th, a, b);
        }
        if (hit && (th < t.bt || (th == t.bt && s < t.bslot))) {
            t.bt = th;
            t.bslot = s;
            t.bu = a;
            t.bv = b;
            if (ANY) return true;
        }
    }
    return false;
}

// Test the primitives of the current leaf, unless a closer hit found since it was pushed culls it.
// Precondition: at_leaf(t).
template <int S, bool ANY, bool COUNT = true>
__device__ __forceinline__ void leaf_step(Trav &t, const StackMem &m, const LeafRecord *__restrict__ rec) {
    if (t.cur_tn <= t.bt * kRobust && test_leaf<ANY, COUNT>(t, t.cur, rec)) {
        t.top = 0;  // any hit: done
        t.nspill = 0;
    }
    t.have = false;
    if (t.top > 0) pop<S>(t, m);
}

// Wave-level choice of which lanes advance this iteration: leaves when enough lanes
// wait on a leaf (or no lane is on a node), nodes otherwise.
__device__ __forceinline__ bool choose_leaf_phase(bool on_node, bool on_leaf, uint32_t leaf_weight) {
    const unsigned long long mn = __ballot(on_node), ml = __ballot(on_leaf);
    if (ml == 0ull) return false;
    if (mn == 0ull) return true;
    return (uint32_t)__popcll(ml) * leaf_weight >= (uint32_t)__popcll(mn);
}

// Run one ray to completion (used where lanes are not refilled).
template <int W, int S, bool ANY, bool COUNT = true>
__device__ __forceinline__ void trav_run(Trav &t, const StackMem &m, const uint32_t *wide, const LeafRecord *rec,
                                         uint32_t leaf_weight) {
    for (;;) {
        const bool on_node = at_node(t), on_leaf = at_leaf(t);
        if (__ballot(on_node || on_leaf) == 0ull) break;
        if (choose_leaf_phase(on_node, on_leaf, leaf_weight)) {
            if (on_leaf) leaf_step<S, ANY, COUNT>(t, m, rec);
        } else {
            if (on_node) node_step<W, S, ANY, COUNT>(t, m, wide);
        }
    }
}

}  // namespace rtgpu
