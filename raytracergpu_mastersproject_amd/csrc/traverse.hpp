// traverse.hpp -- wide-BVH traversal step shared by the ray-query and path-tracing kernels.
//
// One lane owns one ray.  Each call of step() processes one stack entry: a wide
// node (W slab tests, hit children sorted near-to-far, nearest kept as the
// current entry, the rest pushed) or a leaf (its primitives tested).  The
// stack keeps its newest `S` entries in an LDS column private to the lane and
// spills older ones to a private (scratch) array, so deep trees never fault.
//
// Closest-hit semantics are those of hitBVH (raytraceBVH.comp:195-265) for any
// traversal order: the smallest t wins and, among equal t, the smallest leaf
// slot wins -- exactly the primitive the reference's right-first traversal
// tests last, because it accepts equal t (`t > tMax` rejects, :130).
//
// The visit sequence of a ray depends only on the ray and the tree, never on
// scheduling, so per-ray counts of AABB and primitive tests are reproducible.
#pragma once

#include "rt_device.hpp"

namespace rtgpu {

constexpr uint32_t kSpillCap = 184;  // private spill entries per lane (beyond the LDS part)

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

// per-lane traversal state
struct Trav {
    v3 o, d, idir;
    float tmin, bt;     // ray start; best t so far (initially tmax)
    uint32_t bslot;     // leaf slot of the best hit (kEmpty = none)
    float bu, bv;       // barycentrics of the best hit
    uint32_t cur;       // current entry (node index or leaf entry)
    float cur_tn;       // its entry distance
    bool have;          // cur is valid
    uint32_t top, nspill;
    uint32_t nbox, ntri;
};

struct StackMem {
    uint2 *col;       // this lane's LDS column: entry k at col[k * stride]
    uint32_t stride;  // block size
    uint2 *spill;     // private spill array (kSpillCap entries)
};

template <int S> __device__ __forceinline__ void push(Trav &t, const StackMem &m, uint32_t e, float tn) {
    if (t.top - t.nspill == (uint32_t)S) {  // LDS part full: move its oldest entry out
        if (t.nspill < kSpillCap) m.spill[t.nspill] = m.col[(t.nspill & (S - 1)) * m.stride];
        t.nspill++;
    }
    m.col[(t.top & (S - 1)) * m.stride] = make_uint2(e, __float_as_uint(tn));
    t.top++;
}

template <int S> __device__ __forceinline__ void pop(Trav &t, const StackMem &m) {
    uint2 v;
    if (t.top > t.nspill) {
        t.top--;
        v = m.col[(t.top & (S - 1)) * m.stride];
    } else {
        t.nspill--;
        t.top--;
        v = (t.nspill < kSpillCap) ? m.spill[t.nspill] : make_uint2(kEmpty, 0u);
    }
    t.cur = v.x;
    t.cur_tn = __uint_as_float(v.y);
    t.have = v.x != kEmpty;
}

__device__ __forceinline__ void trav_init(Trav &t, v3 o, v3 d, float tmin, float tmax, uint32_t nwide) {
    t.o = o;
    t.d = d;
    t.idir = v3{1.0f / d.x, 1.0f / d.y, 1.0f / d.z};
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

template <int W>
__device__ __forceinline__ void cswap(float *k, uint32_t *ix, uint32_t *e, float *tn, int i, int j) {
    bool sw = (k[i] > k[j]) || (k[i] == k[j] && ix[i] > ix[j]);
    if (sw) {
        float a = k[i]; k[i] = k[j]; k[j] = a;
        uint32_t b = ix[i]; ix[i] = ix[j]; ix[j] = b;
        uint32_t c = e[i]; e[i] = e[j]; e[j] = c;
        float d = tn[i]; tn[i] = tn[j]; tn[j] = d;
    }
}

template <int W> __device__ __forceinline__ void sort_children(float *k, uint32_t *ix, uint32_t *e, float *tn) {
    if constexpr (W == 2) {
        cswap<W>(k, ix, e, tn, 0, 1);
    } else if constexpr (W == 4) {
        cswap<W>(k, ix, e, tn, 0, 1); cswap<W>(k, ix, e, tn, 2, 3);
        cswap<W>(k, ix, e, tn, 0, 2); cswap<W>(k, ix, e, tn, 1, 3);
        cswap<W>(k, ix, e, tn, 1, 2);
    } else {  // Batcher odd-even merge sort, 19 comparators
        cswap<W>(k, ix, e, tn, 0, 1); cswap<W>(k, ix, e, tn, 2, 3); cswap<W>(k, ix, e, tn, 4, 5); cswap<W>(k, ix, e, tn, 6, 7);
        cswap<W>(k, ix, e, tn, 0, 2); cswap<W>(k, ix, e, tn, 1, 3); cswap<W>(k, ix, e, tn, 4, 6); cswap<W>(k, ix, e, tn, 5, 7);
        cswap<W>(k, ix, e, tn, 1, 2); cswap<W>(k, ix, e, tn, 5, 6);
        cswap<W>(k, ix, e, tn, 0, 4); cswap<W>(k, ix, e, tn, 1, 5); cswap<W>(k, ix, e, tn, 2, 6); cswap<W>(k, ix, e, tn, 3, 7);
        cswap<W>(k, ix, e, tn, 2, 4); cswap<W>(k, ix, e, tn, 3, 5);
        cswap<W>(k, ix, e, tn, 1, 2); cswap<W>(k, ix, e, tn, 3, 4); cswap<W>(k, ix, e, tn, 5, 6);
    }
}

// One traversal step.  ANY: stop at the first hit.
template <int W, int S, bool ANY>
__device__ __forceinline__ void trav_step(Trav &t, const StackMem &m, const uint32_t *__restrict__ wide,
                                          const LeafRecord *__restrict__ rec) {
    if (t.have) {
        if (t.cur_tn > t.bt * kRobust) {
            t.have = false;  // culled by a closer hit found since it was pushed
        } else if (t.cur & kLeafBit) {
            const uint32_t first = leaf_first(t.cur), cnt = leaf_count(t.cur);
            for (uint32_t s = first; s < first + cnt; ++s) {
                t.ntri++;
                const LeafRecord r = rec[s];
                float th = 0.0f, a = 0.0f, b = 0.0f;
                bool hit = is_sphere_record(r) ? sphere_test(r, t.o, t.d, t.tmin, t.bt, th)
                                               : triangle_test(r, t.o, t.d, t.tmin, t.bt, th, a, b);
                if (hit && (th < t.bt || (th == t.bt && s < t.bslot))) {
                    t.bt = th;
                    t.bslot = s;
                    t.bu = a;
                    t.bv = b;
                    if (ANY) {
                        t.top = 0;
                        t.nspill = 0;
                        break;
                    }
                }
            }
            t.have = false;
        } else {
            WideNode<W> nd;
            load_node<W>(wide, t.cur, nd);
            float key[W], tnv[W];
            uint32_t ix[W], ent[W];
            int nh = 0;
            const float inf = __uint_as_float(0x7F800000u);
#pragma unroll
            for (int c = 0; c < W; ++c) {
                const bool valid = nd.child[c] != kEmpty;
                t.nbox += valid ? 1u : 0u;
                float x0 = (nd.lo[0][c] - t.o.x) * t.idir.x, x1 = (nd.hi[0][c] - t.o.x) * t.idir.x;
                float y0 = (nd.lo[1][c] - t.o.y) * t.idir.y, y1 = (nd.hi[1][c] - t.o.y) * t.idir.y;
                float z0 = (nd.lo[2][c] - t.o.z) * t.idir.z, z1 = (nd.hi[2][c] - t.o.z) * t.idir.z;
                float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), t.tmin));
                float tf = fminf(fminf(fmaxf(x0, x1), fmaxf(y0, y1)), fminf(fmaxf(z0, z1), t.bt));
                const bool hit = valid && tn <= tf * kRobust;
                nh += hit ? 1 : 0;
                key[c] = hit ? (ANY ? 0.0f : tn) : inf;
                tnv[c] = tn;
                ix[c] = (uint32_t)c;
                ent[c] = nd.child[c];
            }
            sort_children<W>(key, ix, ent, tnv);
            if (nh == 0) {
                t.have = false;
            } else {
#pragma unroll
                for (int k = W - 1; k >= 1; --k)
                    if (k < nh) push<S>(t, m, ent[k], tnv[k]);
                t.cur = ent[0];
                t.cur_tn = tnv[0];
            }
        }
    }
    if (!t.have && t.top > 0) pop<S>(t, m);
}

}  // namespace rtgpu
