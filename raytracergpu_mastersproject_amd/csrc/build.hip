// build.hip -- device BVH construction for a resident scene.
//
// Replaces compute stage S1 of the reference (RaytracerGPU_MastersProject/
// RaytracerBVH.cpp:734-997): model->world transform, centroid bounds, Morton
// codes, a stable radix sort, Karras 2012 topology and a bottom-up refit,
// producing the reference's binary HLBVHNode buffer (the RaytracerBVH builder
// output).  That tree is then collapsed, level by level, into the wide SoA node
// layout the traversal kernels read, with the triangle data re-laid out per
// leaf slot (LeafRecord).  The reference's single-workgroup reduction and sort
// are replaced by device-wide kernels.
#include <algorithm>
#include <cstring>
#include <vector>

#include "rt_internal.hpp"

namespace rtgpu {
namespace {

constexpr int kB = 256;  // build kernels' block size

__device__ inline uint32_t fkey(float f) {
    uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline float fkey_dec(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

inline uint32_t grid_for(uint64_t n, int block = kB) {
    uint64_t g = (n + block - 1) / block;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, 1u << 30));
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ModelSpaceToWorldSpace.comp:32-48: v' = M * vec4(v, 1) evaluated left to right
__device__ inline void mat_vec1(const float *m, float *v) {
    float x = v[0], y = v[1], z = v[2];
    float r[4];
    for (int i = 0; i < 4; ++i) r[i] = m[0 + i] * x + m[4 + i] * y + m[8 + i] * z + m[12 + i] * 1.0f;
    v[0] = r[0]; v[1] = r[1]; v[2] = r[2]; v[3] = r[3];
}
__global__ void k_model_to_world(rt_triangle *tris, uint32_t ntri, rt_sphere *sph, uint32_t nsph,
                                 const float *models) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ntri) {
        rt_triangle t = tris[i];
        const float *m = models + 16u * t.modelIndex;
        mat_vec1(m, t.v0);
        mat_vec1(m, t.v1);
        mat_vec1(m, t.v2);
        tris[i] = t;
    } else if (i < ntri + nsph) {
        rt_sphere s = sph[i - ntri];
        float c[4] = {s.center[0], s.center[1], s.center[2], s.center[3]};
        mat_vec1(models + 16u * s.modelIndex, c);  // radius stays unscaled
        s.center[0] = c[0]; s.center[1] = c[1]; s.center[2] = c[2]; s.center[3] = c[3];
        sph[i - ntri] = s;
    }
}

// centroid of primitive i (GetEnclosingAABB.comp:26-28, GenerateMortonCodesOfPrimitives.comp:62-64)
__device__ inline v3 centroid(const rt_triangle *tris, uint32_t ntri, const rt_sphere *sph, uint32_t i) {
    if (i < ntri) {
        const rt_triangle &t = tris[i];
        v3 s = (v3{t.v0[0], t.v0[1], t.v0[2]} + v3{t.v1[0], t.v1[1], t.v1[2]}) + v3{t.v2[0], t.v2[1], t.v2[2]};
        return s / 3.0f;
    }
    const rt_sphere &s = sph[i - ntri];
    return v3{s.center[0], s.center[1], s.center[2]};
}

__global__ void k_bounds_init(uint32_t *keys) {
    if (threadIdx.x < 3) keys[threadIdx.x] = fkey(1000000000.0f);
    else if (threadIdx.x < 6) keys[threadIdx.x] = fkey(-1000000000.0f);
}

// device-wide min/max of centroids in a total order (-0 < +0): order-independent result
__global__ void k_centroid_bounds(const rt_triangle *tris, uint32_t ntri, const rt_sphere *sph, uint32_t n,
                                  uint32_t *keys) {
    uint32_t mn[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, mx[3] = {0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        v3 c = centroid(tris, ntri, sph, i);
        uint32_t k[3] = {fkey(c.x), fkey(c.y), fkey(c.z)};
        for (int a = 0; a < 3; ++a) {
            mn[a] = min(mn[a], k[a]);
            mx[a] = max(mx[a], k[a]);
        }
    }
    __shared__ uint32_t red[6][kB / 64];
    for (int a = 0; a < 3; ++a) {
        for (int off = 32; off > 0; off >>= 1) {
            mn[a] = min(mn[a], (uint32_t)__shfl_xor((int)mn[a], off));
            mx[a] = max(mx[a], (uint32_t)__shfl_xor((int)mx[a], off));
        }
    }
    int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0)
        for (int a = 0; a < 3; ++a) { red[a][wave] = mn[a]; red[3 + a][wave] = mx[a]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / 64); ++w)
            for (int a = 0; a < 3; ++a) {
                red[a][0] = min(red[a][0], red[a][w]);
                red[3 + a][0] = max(red[3 + a][0], red[3 + a][w]);
            }
        for (int a = 0; a < 3; ++a) {
            atomicMin(&keys[a], red[a][0]);
            atomicMax(&keys[3 + a], red[3 + a][0]);
        }
    }
}

// padAABB (GetEnclosingAABB.comp:33-46); out: emin[3], emax[3], span[3]
__global__ void k_bounds_finalize(const uint32_t *keys, float *bounds) {
    if (threadIdx.x != 0) return;
    const float delta = 0.001f, pad = 0.001f / 2.0f;
    for (int a = 0; a < 3; ++a) {
        float mn = fkey_dec(keys[a]), mx = fkey_dec(keys[3 + a]);
        if (mx - mn < delta) { mn -= pad; mx += pad; }
        bounds[a] = mn;
        bounds[3 + a] = mx;
        bounds[6 + a] = mx - mn;
    }
}

__device__ inline uint32_t sep3(uint32_t v) {  // GenerateMortonCodesOfPrimitives.comp:20-28
    if (v == 1024u) v--;
    v = (v | (v << 16)) & 50331903u;
    v = (v | (v << 8)) & 50393103u;
    v = (v | (v << 4)) & 51130563u;
    v = (v | (v << 2)) & 153391689u;
    return v;
}

__global__ void k_morton(const rt_triangle *tris, uint32_t ntri, const rt_sphere *sph, uint32_t n,
                         const float *bounds, uint32_t ref_morton, uint32_t *keys, uint32_t *vals) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    v3 c = centroid(tris, ntri, sph, i);
    float cc[3] = {c.x, c.y, c.z};
    uint32_t q[3];
    for (int a = 0; a < 3; ++a) {
        float off = ref_morton ? cc[a] / bounds[6 + a] : (cc[a] - bounds[a]) / bounds[6 + a];
        q[a] = f2u_sat(off * 1024.0f);
    }
    keys[i] = (sep3(q[2]) << 2) | (sep3(q[1]) << 1) | sep3(q[0]);
    vals[i] = i;
}

// ---------------------------------------------------------------------------
// Exclusive scan of u32 (device-wide, recursive over 2048-element blocks)
// ---------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kB * kScanItems;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t wsum[kB / 64];
    int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    uint32_t x = v;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = (uint32_t)__shfl_up((int)x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t base = 0, all = 0;
    for (int w = 0; w < kB / 64; ++w) {
        if (w < wave) base += wsum[w];
        all += wsum[w];
    }
    __syncthreads();
    *total = all;
    return base + x - v;
}

__global__ void k_scan_block(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *block_sums) {
    uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems], s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t pre = block_exclusive_scan(s, &tot);
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) out[base + k] = pre;
        pre += v[k];
    }
    if (threadIdx.x == 0 && block_sums) block_sums[blockIdx.x] = tot;
}
__global__ void k_scan_add(uint32_t *out, uint32_t n, const uint32_t *block_offsets) {
    uint64_t i = (uint64_t)blockIdx.x * kScanTile + threadIdx.x;
    uint32_t add = block_offsets[blockIdx.x];
    for (int k = 0; k < kScanItems; ++k, i += kB)
        if (i < n) out[i] += add;
}
__global__ void k_scan_total(const uint32_t *in, const uint32_t *out, uint32_t n, uint32_t *total) {
    if (threadIdx.x == 0) *total = n ? out[n - 1] + in[n - 1] : 0u;
}

// Temporaries of one build, carved from one block of the context's scratch cache.  Every
// launch of a build is on one stream, so a range can be handed out again as soon as the
// launches using it have been enqueued (mark / release); the block itself returns to the
// cache with an event after the build's last launch.
class Temps {
  public:
    Temps(rt_context *ctx, uint64_t bytes, hipStream_t st) : lease_(ctx->scratch, bytes, st), cap_(bytes) {}
    template <class T> T *take(uint64_t count) {
        const uint64_t b = (sizeof(T) * count + 255u) & ~uint64_t(255);
        RT_REQUIRE(used_ + b <= cap_, "internal: build temporaries exceed their bound");
        T *p = reinterpret_cast<T *>(lease_.ptr() + used_);
        used_ += b;
        return p;
    }
    uint64_t mark() const { return used_; }
    void release(uint64_t m) { used_ = m; }

  private:
    ScratchLease lease_;
    uint64_t cap_, used_ = 0;
};

uint64_t round256(uint64_t b) { return (b + 255u) & ~uint64_t(255); }

// bytes exclusive_scan takes from Temps for n elements
uint64_t scan_temp_bytes(uint64_t n) {
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    return nb <= 1 ? 0 : 2 * round256(4 * nb) + scan_temp_bytes(nb);
}

void exclusive_scan(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *d_total, hipStream_t st, Temps &tmp) {
    uint32_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb <= 1) {
        k_scan_block<<<1, kB, 0, st>>>(in, out, n, nullptr);
    } else {
        const uint64_t m = tmp.mark();
        uint32_t *sums = tmp.take<uint32_t>(nb), *sums_scan = tmp.take<uint32_t>(nb);
        k_scan_block<<<nb, kB, 0, st>>>(in, out, n, sums);
        exclusive_scan(sums, sums_scan, nb, nullptr, st, tmp);
        k_scan_add<<<nb, kB, 0, st>>>(out, n, sums_scan);
        tmp.release(m);
    }
    if (d_total) k_scan_total<<<1, 64, 0, st>>>(in, out, n, d_total);
}

// ---------------------------------------------------------------------------
// Stable LSD radix sort, 4 passes x 8 bits (RadixSortSimple.comp's intent, A14):
// per-tile digit histograms -> digit-major exclusive scan -> stable scatter in
// which each round ranks 256 keys with wave ballots (keys keep tile order).
// ---------------------------------------------------------------------------
constexpr int kSortItems = 8;
constexpr int kSortTile = kB * kSortItems;

__global__ void k_radix_hist(const uint32_t *keys, uint32_t n, uint32_t shift, uint32_t *ghist, uint32_t nblocks) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        uint64_t p = base + (uint64_t)r * kB + threadIdx.x;
        if (p < n) atomicAdd(&h[(keys[p] >> shift) & 255u], 1u);
    }
    __syncthreads();
    ghist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ void k_radix_scatter(const uint32_t *keys_in, const uint32_t *vals_in, uint32_t *keys_out,
                                uint32_t *vals_out, uint32_t n, uint32_t shift, const uint32_t *goff,
                                uint32_t nblocks) {
    __shared__ uint32_t off[256];
    __shared__ uint32_t cnt[kB / 64][256];
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    off[tid] = goff[(uint64_t)tid * nblocks + blockIdx.x];
    const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
    uint64_t base = (uint64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        for (int w = 0; w < kB / 64; ++w) cnt[w][tid] = 0;
        __syncthreads();
        uint64_t p = base + (uint64_t)r * kB + tid;
        bool valid = p < n;
        uint32_t key = valid ? keys_in[p] : 0u, val = valid ? vals_in[p] : 0u;
        uint32_t digit = (key >> shift) & 255u;
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            unsigned long long bb = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? bb : ~bb;
        }
        uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) cnt[wave][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t before = 0;
            for (int w = 0; w < wave; ++w) before += cnt[w][digit];
            uint32_t dst = off[digit] + before + rank;
            keys_out[dst] = key;
            vals_out[dst] = val;
        }
        __syncthreads();
        uint32_t add = 0;
        for (int w = 0; w < kB / 64; ++w) add += cnt[w][tid];
        off[tid] += add;
        __syncthreads();
    }
}

uint64_t radix_temp_bytes(uint64_t n) {
    const uint64_t nblocks = (n + kSortTile - 1) / kSortTile;
    return 2 * round256(4 * n) + 2 * round256(4 * 256 * nblocks) + scan_temp_bytes(256 * nblocks);
}

void radix_sort(uint32_t *keys, uint32_t *vals, uint32_t n, hipStream_t st, Temps &tmp) {
    if (n <= 1) return;
    uint32_t nblocks = (n + kSortTile - 1) / kSortTile;
    const uint64_t m = tmp.mark();
    uint32_t *k2 = tmp.take<uint32_t>(n), *v2 = tmp.take<uint32_t>(n);
    uint32_t *hist = tmp.take<uint32_t>(256ull * nblocks), *hoff = tmp.take<uint32_t>(256ull * nblocks);
    uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
    for (int pass = 0; pass < 4; ++pass) {  // an even number of passes: the result is back in keys/vals
        uint32_t shift = 8u * pass;
        k_radix_hist<<<nblocks, kB, 0, st>>>(ka, n, shift, hist, nblocks);
        exclusive_scan(hist, hoff, 256u * nblocks, nullptr, st, tmp);
        k_radix_scatter<<<nblocks, kB, 0, st>>>(ka, va, kb, vb, n, shift, hoff, nblocks);
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    tmp.release(m);
}

// ---------------------------------------------------------------------------
// Karras topology (ConstructHLBVH.comp:58-106, 171-214)
// ---------------------------------------------------------------------------
__device__ inline int kdelta(const uint32_t *codes, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    uint32_t ci = codes[i], cj = codes[j];
    if (ci == cj) return 32 + __clz((int)((uint32_t)i ^ (uint32_t)j));
    return __clz((int)(ci ^ cj));
}

__global__ void k_karras(const uint32_t *codes, int n, rt_bvh_node *nodes, uint32_t *parent, uint2 *range) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id == 0) parent[0] = 0;
    if (id >= n - 1) return;
    int dl = kdelta(codes, n, id, id - 1), dr = kdelta(codes, n, id, id + 1);
    int dir = (dr >= dl) ? 1 : -1;
    int dmin = min(dl, dr);
    int lmax = 2;
    while (kdelta(codes, n, id, id + lmax * dir) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t > 0; t >>= 1)
        if (kdelta(codes, n, id, id + (l + t) * dir) > dmin) l += t;
    int e = id + l * dir;
    int first = min(id, e), last = max(id, e);
    int common = kdelta(codes, n, first, last);
    int split = first, stride = last - first;
    do {
        stride = (stride + 1) >> 1;
        int ns = split + stride;
        if (ns < last && kdelta(codes, n, first, ns) > common) split = ns;
    } while (stride > 1);
    const int leaf0 = n - 1;
    int left = (split == first) ? leaf0 + split : split;
    int right = (split + 1 == last) ? leaf0 + split + 1 : split + 1;
    rt_bvh_node nd;
    nd.minX = nd.maxX = nd.minY = nd.maxY = nd.minZ = nd.maxZ = 0.0f;
    nd.left = (uint32_t)left;
    nd.right = (uint32_t)right;
    nd.primitiveIndex = 0;
    nd.primitiveType = 0;
    nodes[id] = nd;
    parent[left] = (uint32_t)id;
    parent[right] = (uint32_t)id;
    range[id] = make_uint2((uint32_t)first, (uint32_t)last);
}

// leaf slot -> primitive (A11 selects the reference's unsorted placement)
__global__ void k_slots(const uint32_t *sorted_vals, uint32_t n, uint32_t ntri, uint32_t ref_leaves,
                        uint32_t *slot_prim) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t p = ref_leaves ? i : sorted_vals[i];
    slot_prim[i] = p < ntri ? p : (RT_SPHERE_BIT | (p - ntri));
}

// leaf boxes (ConstructHLBVH.comp:33-56, 134-169), leaf records and material table
__device__ inline void leaf_body(const rt_triangle *tris, const rt_sphere *sph, const float2 *sph_uv,
                                 const uint32_t *slot_prim, uint32_t i, rt_bvh_node *leaf_nodes, LeafRecord *records,
                                 uint32_t *slot_mat) {
    uint32_t p = slot_prim[i];
    rt_bvh_node nd;
    LeafRecord rec;
    if (p & RT_SPHERE_BIT) {
        const rt_sphere &s = sph[p & ~RT_SPHERE_BIT];
        float l[3], r[3];
        for (int a = 0; a < 3; ++a) { l[a] = s.center[a] - s.radius; r[a] = s.center[a] + s.radius; }
        nd.minX = gmin(l[0], r[0]); nd.maxX = gmax(l[0], r[0]);
        nd.minY = gmin(l[1], r[1]); nd.maxY = gmax(l[1], r[1]);
        nd.minZ = gmin(l[2], r[2]); nd.maxZ = gmax(l[2], r[2]);
        nd.primitiveIndex = p & ~RT_SPHERE_BIT;
        nd.primitiveType = RT_SPHERE_PRIMITIVE;
        rec.r0 = make_float4(s.center[0], s.center[1], s.center[2], s.radius);
        const float2 uv = sph_uv[p & ~RT_SPHERE_BIT];
        rec.r1 = make_float4(__uint_as_float(kSphereTag), uv.x, uv.y, 0.0f);
        rec.r2 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        rec.r3 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        slot_mat[i] = s.materialIndex;
    } else {
        const rt_triangle &t = tris[p];
        nd.minX = gmin(t.v0[0], gmin(t.v1[0], t.v2[0]));
        nd.maxX = gmax(t.v0[0], gmax(t.v1[0], t.v2[0]));
        nd.minY = gmin(t.v0[1], gmin(t.v1[1], t.v2[1]));
        nd.maxY = gmax(t.v0[1], gmax(t.v1[1], t.v2[1]));
        nd.minZ = gmin(t.v0[2], gmin(t.v1[2], t.v2[2]));
        nd.maxZ = gmax(t.v0[2], gmax(t.v1[2], t.v2[2]));
        nd.primitiveIndex = p;
        nd.primitiveType = RT_TRIANGLE_PRIMITIVE;
        rec = make_triangle_record(v3{t.v0[0], t.v0[1], t.v0[2]}, v3{t.v1[0], t.v1[1], t.v1[2]},
                                   v3{t.v2[0], t.v2[1], t.v2[2]});
        slot_mat[i] = t.materialIndex;
    }
    const float delta = 0.001f, pad = 0.001f / 2.0f;
    if (nd.maxX - nd.minX < delta) { nd.minX -= pad; nd.maxX += pad; }
    if (nd.maxY - nd.minY < delta) { nd.minY -= pad; nd.maxY += pad; }
    if (nd.maxZ - nd.minZ < delta) { nd.minZ -= pad; nd.maxZ += pad; }
    nd.left = 0;
    nd.right = 0;
    if (leaf_nodes) leaf_nodes[i] = nd;
    records[i] = rec;
}
__global__ void k_leaves(const rt_triangle *tris, const rt_sphere *sph, const float2 *sph_uv, const uint32_t *slot_prim,
                         uint32_t n, rt_bvh_node *leaf_nodes, LeafRecord *records, uint32_t *slot_mat) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) leaf_body(tris, sph, sph_uv, slot_prim, i, leaf_nodes, records, slot_mat);
}
// leaf records only (caller-supplied node buffer: boxes come from that buffer)
__global__ void k_records(const rt_triangle *tris, const rt_sphere *sph, const float2 *sph_uv, const uint32_t *slot_prim,
                          uint32_t n, LeafRecord *records, uint32_t *slot_mat) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) leaf_body(tris, sph, sph_uv, slot_prim, i, nullptr, records, slot_mat);
}

// ConstructAABBsOfInternalNodes.comp:27-65 computes each internal box as the union of
// its children's, bottom-up, with an atomic visit counter per node (the second visitor
// writes the union).  Here the same unions are computed level by level instead, deepest
// level first, one launch per level: on this device the eight XCDs keep separate L2
// caches, so each of the counter scheme's release/acquire pairs writes back and
// invalidates a whole L2, while launch boundaries make a level's boxes visible for free.
// Depths come from pointer jumping over the parent links.  In a Karras tree every child's
// range shares a strictly longer prefix of the 64-bit (code, index) key than its parent's
// (kdelta below), so the depth is at most 64 and the per-level launches stay few.

// internal node i: jump = parent (the root points at itself), dist = 1 (root 0)
__global__ void k_depth_init(const uint32_t *parent, uint32_t m, uint32_t *jump, uint32_t *dist) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    jump[i] = i == 0 ? 0u : parent[i];
    dist[i] = i == 0 ? 0u : 1u;
}

// one doubling step: dist(i) += dist(jump(i)), jump(i) = jump(jump(i)); *more = 1 (one
// store per wave) while a jump has not reached the root
__global__ void k_depth_jump(const uint32_t *jump, const uint32_t *dist, uint32_t *jump2, uint32_t *dist2, uint32_t m,
                             uint32_t *more) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint32_t j = jump[i];
    dist2[i] = dist[i] + dist[j];
    const uint32_t jj = jump[j];
    jump2[i] = jj;
    const unsigned long long pending = __ballot(jj != 0u);
    if (pending != 0ull && __lane_id() == (uint32_t)__ffsll((long long)pending) - 1u) *more = 1u;
}

// largest depth: block maximum in LDS, one global atomic per block
__global__ void k_depth_max(const uint32_t *depth, uint32_t m, uint32_t *out) {
    __shared__ uint32_t best;
    if (threadIdx.x == 0) best = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) atomicMax(&best, depth[i]);
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, best);
}

// the union for every node at depth d (its children are leaves or deeper nodes)
__global__ void k_refit_level(rt_bvh_node *nodes, const uint32_t *depth, uint32_t m, uint32_t d) {
    const uint32_t id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= m || depth[id] != d) return;
    rt_bvh_node nd = nodes[id];
    const rt_bvh_node a = nodes[nd.left], b = nodes[nd.right];
    nd.minX = gmin(a.minX, b.minX); nd.maxX = gmax(a.maxX, b.maxX);
    nd.minY = gmin(a.minY, b.minY); nd.maxY = gmax(a.maxY, b.maxY);
    nd.minZ = gmin(a.minZ, b.minZ); nd.maxZ = gmax(a.maxZ, b.maxZ);
    nodes[id] = nd;
}

// Internal boxes of the n-1 internal nodes (leaf boxes are already written).
uint64_t refit_temp_bytes(uint64_t n) { return 4 * round256(4 * n) + round256(4); }

void refit_levels(rt_bvh_node *nodes, const uint32_t *parent, uint32_t n, hipStream_t st, Temps &tmp) {
    const uint32_t m = n - 1;
    const uint64_t mk = tmp.mark();
    uint32_t *jump = tmp.take<uint32_t>(m), *dist = tmp.take<uint32_t>(m), *jump2 = tmp.take<uint32_t>(m),
             *dist2 = tmp.take<uint32_t>(m), *flag = tmp.take<uint32_t>(1);
    k_depth_init<<<grid_for(m), kB, 0, st>>>(parent, m, jump, dist);
    for (uint32_t round = 0; round < 32; ++round) {  // 2^32 exceeds any depth
        k_fill_u32<<<1, kB, 0, st>>>(flag, 0u, 1u);
        k_depth_jump<<<grid_for(m), kB, 0, st>>>(jump, dist, jump2, dist2, m, flag);
        std::swap(jump, jump2);
        std::swap(dist, dist2);
        uint32_t more = 0;
        RT_HIP(hipMemcpyAsync(&more, flag, sizeof more, hipMemcpyDeviceToHost, st));
        RT_HIP(hipStreamSynchronize(st));
        if (!more) break;
    }
    k_fill_u32<<<1, kB, 0, st>>>(flag, 0u, 1u);
    k_depth_max<<<grid_for(m), kB, 0, st>>>(dist, m, flag);
    uint32_t deepest = 0;
    RT_HIP(hipMemcpyAsync(&deepest, flag, sizeof deepest, hipMemcpyDeviceToHost, st));
    RT_HIP(hipStreamSynchronize(st));
    RT_REQUIRE(deepest < m, "internal: LBVH parent links do not form a tree");
    for (uint32_t d = deepest + 1; d-- > 0;) k_refit_level<<<grid_for(m), kB, 0, st>>>(nodes, dist, m, d);
    tmp.release(mk);
}

// ---------------------------------------------------------------------------
// Collapse binary -> wide, one BFS level per launch pair.  A wide node made
// from binary node b takes b's children and repeatedly opens the child with the
// largest surface area until W children exist; a child whose subtree holds at
// most leaf_size primitives becomes a leaf over its contiguous slot range.
// ---------------------------------------------------------------------------
struct CollapseArgs {
    const rt_bvh_node *nodes;
    const uint2 *range;  // internal nodes: leaf-slot range of the subtree
    uint32_t n;          // primitives
    uint32_t width, leaf_size;
};

__device__ inline uint32_t subtree_count(const CollapseArgs &a, uint32_t b) {
    if (b >= a.n - 1 || (a.nodes[b].left == 0 && a.nodes[b].right == 0)) return 1u;
    uint2 r = a.range[b];
    return r.y - r.x + 1u;
}
__device__ inline bool is_internal(const CollapseArgs &a, uint32_t b) {
    return a.n > 1 && b < a.n - 1;
}
__device__ inline float box_area(const rt_bvh_node &nd) {
    float dx = nd.maxX - nd.minX, dy = nd.maxY - nd.minY, dz = nd.maxZ - nd.minZ;
    return dx * dy + dy * dz + dz * dx;
}

__device__ int choose_children(const CollapseArgs &a, uint32_t b, uint32_t *cand) {
    int nc;
    if (!is_internal(a, b)) {
        cand[0] = b;
        return 1;
    }
    cand[0] = a.nodes[b].left;
    cand[1] = a.nodes[b].right;
    nc = 2;
    while (nc < (int)a.width) {
        int best = -1;
        float best_area = -1.0f;
        for (int k = 0; k < nc; ++k) {
            uint32_t c = cand[k];
            if (!is_internal(a, c) || subtree_count(a, c) <= a.leaf_size) continue;
            float ar = box_area(a.nodes[c]);
            if (ar > best_area) { best_area = ar; best = k; }
        }
        if (best < 0) break;
        uint32_t c = cand[best];
        for (int k = nc; k > best + 1; --k) cand[k] = cand[k - 1];
        cand[best] = a.nodes[c].left;
        cand[best + 1] = a.nodes[c].right;
        ++nc;
    }
    return nc;
}

__device__ inline bool becomes_node(const CollapseArgs &a, uint32_t c) {
    return is_internal(a, c) && subtree_count(a, c) > a.leaf_size;
}

__global__ void k_collapse_count(CollapseArgs a, const uint32_t *queue, uint32_t qn, uint32_t *newcount) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= qn) return;
    uint32_t cand[8];
    int nc = choose_children(a, queue[e], cand);
    uint32_t k = 0;
    for (int i = 0; i < nc; ++i) k += becomes_node(a, cand[i]) ? 1u : 0u;
    newcount[e] = k;
}

__global__ void k_collapse_emit(CollapseArgs a, const uint32_t *queue, uint32_t qn, uint32_t level_base,
                                uint32_t next_base, const uint32_t *child_base, uint32_t *next_queue,
                                uint32_t *wide) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= qn) return;
    const uint32_t W = a.width;
    uint32_t cand[8];
    int nc = choose_children(a, queue[e], cand);
    uint32_t *nd = wide + (size_t)(level_base + e) * 8u * W;
    float *f = reinterpret_cast<float *>(nd);
    uint32_t j = child_base[e];
    const float inf = __uint_as_float(0x7F800000u);
    for (uint32_t k = 0; k < W; ++k) {
        if ((int)k >= nc) {
            f[0 * W + k] = inf; f[1 * W + k] = inf;
            f[2 * W + k] = inf; f[3 * W + k] = inf;
            f[4 * W + k] = inf; f[5 * W + k] = inf;
            nd[6 * W + k] = kEmpty;
            nd[7 * W + k] = 0u;
            continue;
        }
        uint32_t c = cand[k];
        const rt_bvh_node bn = a.nodes[c];
        f[0 * W + k] = bn.minX; f[1 * W + k] = bn.maxX;
        f[2 * W + k] = bn.minY; f[3 * W + k] = bn.maxY;
        f[4 * W + k] = bn.minZ; f[5 * W + k] = bn.maxZ;
        uint32_t ce;
        if (becomes_node(a, c)) {
            ce = next_base + j;
            next_queue[j] = c;
            ++j;
        } else if (is_internal(a, c)) {
            uint2 r = a.range[c];
            ce = leaf_entry(r.x, r.y - r.x + 1u);
        } else {
            ce = leaf_entry(c - (a.n - 1u), 1u);
        }
        nd[6 * W + k] = ce;
        nd[7 * W + k] = 0u;
    }
}

template <class T> void upload(DevBuf &b, const T *src, size_t count) {
    b.alloc(sizeof(T) * count);
    if (count) RT_HIP(hipMemcpy(b.p, src, sizeof(T) * count, hipMemcpyHostToDevice));
}

// Validate a caller-supplied binary node buffer and derive leaf slots and subtree ranges.
void validate_nodes(const rt_bvh_node *nodes, uint32_t n, uint32_t ntri, uint32_t nsph,
                    std::vector<uint32_t> &slot_prim, std::vector<uint2> &range, uint32_t &leaf_size) {
    slot_prim.assign(n, 0);
    range.assign(n > 1 ? n - 1 : 1, make_uint2(0, 0));
    std::vector<uint8_t> seen(2u * n - 1u, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const rt_bvh_node &lf = nodes[n - 1 + i];
        RT_REQUIRE(lf.left == 0 && lf.right == 0, "node buffer: node N-1+i must be a leaf");
        if (lf.primitiveType == RT_TRIANGLE_PRIMITIVE) {
            RT_REQUIRE(lf.primitiveIndex < ntri, "node buffer: triangle index out of range");
            slot_prim[i] = lf.primitiveIndex;
        } else {
            RT_REQUIRE(lf.primitiveType == RT_SPHERE_PRIMITIVE && lf.primitiveIndex < nsph,
                       "node buffer: bad sphere leaf");
            slot_prim[i] = RT_SPHERE_BIT | lf.primitiveIndex;
        }
    }
    if (n == 1) return;
    // iterative post-order over internal nodes: reachability, no cycles, ranges
    std::vector<uint32_t> cnt(n - 1, 0);
    std::vector<uint32_t> stack;
    stack.push_back(0);
    seen[0] = 1;
    bool contiguous = true;
    std::vector<uint8_t> expanded(n - 1, 0);
    while (!stack.empty()) {
        uint32_t id = stack.back();
        const rt_bvh_node &nd = nodes[id];
        RT_REQUIRE(nd.left < 2u * n - 1u && nd.right < 2u * n - 1u && nd.left != 0 && nd.right != 0,
                   "node buffer: child index out of range");
        if (!expanded[id]) {
            expanded[id] = 1;
            for (uint32_t c : {nd.right, nd.left}) {
                RT_REQUIRE(!seen[c], "node buffer: node reachable twice (cycle or shared child)");
                seen[c] = 1;
                if (c < n - 1) stack.push_back(c);
            }
            continue;
        }
        stack.pop_back();
        uint32_t lo = 0xFFFFFFFFu, hi = 0, total = 0;
        for (uint32_t c : {nd.left, nd.right}) {
            if (c >= n - 1) { lo = std::min(lo, c - (n - 1)); hi = std::max(hi, c - (n - 1)); total += 1; }
            else { lo = std::min(lo, range[c].x); hi = std::max(hi, range[c].y); total += cnt[c]; }
        }
        range[id] = make_uint2(lo, hi);
        cnt[id] = total;
        if (hi - lo + 1 != total) contiguous = false;
    }
    for (uint32_t i = 0; i < 2u * n - 1u; ++i) RT_REQUIRE(seen[i], "node buffer: unreachable node");
    if (!contiguous) leaf_size = 1;  // multi-primitive leaves need contiguous slot ranges
}

}  // namespace

void build_scene(rt_context *ctx, Scene &sc, const rt_triangle *tris, uint32_t ntri, const rt_sphere *sph,
                 uint32_t nsph, const rt_material *mats, uint32_t nmat, const float *models, uint32_t nmodels,
                 const rt_bvh_node *nodes, uint32_t nnodes, const rt_build_options &opts) {
    (void)ctx;
    const uint32_t W = opts.width ? opts.width : 4u;
    RT_REQUIRE(W == 2 || W == 4 || W == 8, "width must be 2, 4 or 8");
    uint32_t L = opts.leaf_size ? opts.leaf_size : 1u;
    RT_REQUIRE(L >= 1 && L <= 8, "leaf_size must be 1..8");
    const uint64_t n64 = (uint64_t)ntri + nsph;
    RT_REQUIRE(n64 < (1ull << 27) - 16, "too many primitives (max 2^27-17)");
    const uint32_t n = (uint32_t)n64;
    RT_REQUIRE(ntri == 0 || tris, "tris is NULL");
    RT_REQUIRE(nsph == 0 || sph, "spheres is NULL");
    RT_REQUIRE(nmat == 0 || mats, "materials is NULL");
    RT_REQUIRE(nmodels == 0 || models, "models is NULL");
    for (uint32_t i = 0; i < ntri; ++i) {
        RT_REQUIRE(tris[i].materialIndex < nmat, "triangle " + std::to_string(i) + ": materialIndex out of range");
        RT_REQUIRE(!models || tris[i].modelIndex < nmodels,
                   "triangle " + std::to_string(i) + ": modelIndex out of range");
    }
    for (uint32_t i = 0; i < nsph; ++i) {
        RT_REQUIRE(sph[i].materialIndex < nmat, "sphere " + std::to_string(i) + ": materialIndex out of range");
        RT_REQUIRE(!models || sph[i].modelIndex < nmodels, "sphere " + std::to_string(i) + ": modelIndex out of range");
    }
    if (nodes) RT_REQUIRE(n > 0 && nnodes == 2u * n - 1u, "node buffer must hold 2N-1 nodes");

    sc.ntri = ntri; sc.nsph = nsph; sc.nmat = nmat; sc.nprim = n;
    sc.width = W; sc.flags = opts.flags;
    upload(sc.materials, mats, nmat);
    if (n == 0) {
        sc.nwide = 0; sc.wide_depth = 0; sc.stack_bound = 0; sc.leaf_size = L;
        return;
    }
    hipStream_t st = nullptr;
    hipEvent_t ev0, ev1;
    RT_HIP(hipEventCreate(&ev0));
    RT_HIP(hipEventCreate(&ev1));
    DevBuf dtris, dsph, dmodels;
    upload(dtris, tris, ntri);
    upload(dsph, sph, nsph);
    RT_HIP(hipEventRecord(ev0, st));
    if (models) {
        upload(dmodels, models, 16ull * nmodels);
        k_model_to_world<<<grid_for(n), kB, 0, st>>>(dtris.as<rt_triangle>(), ntri, dsph.as<rt_sphere>(), nsph,
                                                     dmodels.as<float>());
    }
    // centre-based u, v of every sphere (raytraceBVH.comp:172-173), from its world-space centre
    DevBuf duv;
    if (nsph) {
        std::vector<rt_sphere> world(sph, sph + nsph);
        if (models) RT_HIP(hipMemcpy(world.data(), dsph.p, sizeof(rt_sphere) * nsph, hipMemcpyDeviceToHost));
        std::vector<float2> uv(nsph);
        for (uint32_t i = 0; i < nsph; ++i) sphere_uv(world[i].center, uv[i].x, uv[i].y);
        upload(duv, uv.data(), nsph);
    }
    sc.slot_prim.alloc(sizeof(uint32_t) * n);
    sc.slot_mat.alloc(sizeof(uint32_t) * n);
    sc.records.alloc(sizeof(LeafRecord) * n);
    // every temporary of the build from one block of the scratch cache
    const bool keep_binary = (opts.flags & RT_BUILD_KEEP_BINARY) != 0;
    const uint32_t max_nodes = std::max<uint32_t>(n - 1, 1);
    const uint64_t temp_bytes =
        (keep_binary ? 0 : round256(sizeof(rt_bvh_node) * (2ull * n - 1))) + round256(sizeof(uint2) * max_nodes) +
        round256(24) + round256(36) + 2 * round256(4ull * n) + round256(4ull * (2ull * n - 1)) +
        std::max(radix_temp_bytes(n), refit_temp_bytes(n)) +                      // sort, then refit
        round256(32ull * W * max_nodes) + 4 * round256(4ull * max_nodes) + round256(4) + scan_temp_bytes(max_nodes);
    Temps tmp(ctx, temp_bytes, st);
    DevBuf kept;  // the binary nodes, when they stay resident
    rt_bvh_node *binary = nullptr;
    if (keep_binary) {
        kept.alloc(sizeof(rt_bvh_node) * (2ull * n - 1));
        binary = kept.as<rt_bvh_node>();
    } else {
        binary = tmp.take<rt_bvh_node>(2ull * n - 1);
    }
    uint2 *range = tmp.take<uint2>(max_nodes);
    if (nodes) {
        std::vector<uint32_t> hslots;
        std::vector<uint2> hrange;
        validate_nodes(nodes, n, ntri, nsph, hslots, hrange, L);
        RT_HIP(hipMemcpy(binary, nodes, sizeof(rt_bvh_node) * (2ull * n - 1), hipMemcpyHostToDevice));
        RT_HIP(hipMemcpy(sc.slot_prim.p, hslots.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
        RT_HIP(hipMemcpy(range, hrange.data(), sizeof(uint2) * hrange.size(), hipMemcpyHostToDevice));
        k_records<<<grid_for(n), kB, 0, st>>>(dtris.as<rt_triangle>(), dsph.as<rt_sphere>(), duv.as<float2>(),
                                              sc.slot_prim.as<uint32_t>(), n, sc.records.as<LeafRecord>(),
                                              sc.slot_mat.as<uint32_t>());
    } else {
        const uint64_t mk = tmp.mark();
        uint32_t *bkeys = tmp.take<uint32_t>(6);
        float *bounds = tmp.take<float>(9);
        uint32_t *keys = tmp.take<uint32_t>(n), *vals = tmp.take<uint32_t>(n), *parent = tmp.take<uint32_t>(2ull * n - 1);
        k_bounds_init<<<1, 64, 0, st>>>(bkeys);
        uint32_t g = std::min<uint32_t>(grid_for(n), 4096);
        k_centroid_bounds<<<g, kB, 0, st>>>(dtris.as<rt_triangle>(), ntri, dsph.as<rt_sphere>(), n, bkeys);
        k_bounds_finalize<<<1, 64, 0, st>>>(bkeys, bounds);
        k_morton<<<grid_for(n), kB, 0, st>>>(dtris.as<rt_triangle>(), ntri, dsph.as<rt_sphere>(), n, bounds,
                                             (opts.flags & RT_BUILD_REFERENCE_MORTON) ? 1u : 0u, keys, vals);
        radix_sort(keys, vals, n, st, tmp);
        if (n > 1) k_karras<<<grid_for(n - 1), kB, 0, st>>>(keys, (int)n, binary, parent, range);
        k_slots<<<grid_for(n), kB, 0, st>>>(vals, n, ntri, (opts.flags & RT_BUILD_REFERENCE_LEAVES) ? 1u : 0u,
                                            sc.slot_prim.as<uint32_t>());
        k_leaves<<<grid_for(n), kB, 0, st>>>(dtris.as<rt_triangle>(), dsph.as<rt_sphere>(), duv.as<float2>(),
                                             sc.slot_prim.as<uint32_t>(), n, binary + (n - 1),
                                             sc.records.as<LeafRecord>(), sc.slot_mat.as<uint32_t>());
        if (n > 1) refit_levels(binary, parent, n, st, tmp);
        tmp.release(mk);
    }
    RT_HIP(hipGetLastError());

    // collapse into the wide layout, BFS order
    CollapseArgs a{binary, range, n, W, L};
    uint32_t *wide_tmp = tmp.take<uint32_t>(8ull * W * max_nodes);
    uint32_t *qa = tmp.take<uint32_t>(max_nodes), *qb = tmp.take<uint32_t>(max_nodes);
    uint32_t *newcount = tmp.take<uint32_t>(max_nodes), *child_base = tmp.take<uint32_t>(max_nodes);
    uint32_t *d_total = tmp.take<uint32_t>(1);
    k_fill_u32<<<1, 64, 0, st>>>(qa, 0u, 1u);  // the root
    uint32_t qn = 1, level_base = 0, depth = 0;
    uint32_t *qcur = qa, *qnext = qb;
    while (qn > 0) {
        k_collapse_count<<<grid_for(qn), kB, 0, st>>>(a, qcur, qn, newcount);
        exclusive_scan(newcount, child_base, qn, d_total, st, tmp);
        uint32_t next_n = 0;
        RT_HIP(hipMemcpy(&next_n, d_total, sizeof(uint32_t), hipMemcpyDeviceToHost));
        RT_REQUIRE(level_base + qn + next_n <= max_nodes, "internal: wide node count exceeds bound");
        k_collapse_emit<<<grid_for(qn), kB, 0, st>>>(a, qcur, qn, level_base, level_base + qn, child_base, qnext,
                                                     wide_tmp);
        level_base += qn;
        qn = next_n;
        std::swap(qcur, qnext);
        ++depth;
    }
    RT_HIP(hipEventRecord(ev1, st));
    RT_HIP(hipEventSynchronize(ev1));
    RT_HIP(hipGetLastError());
    float ms = 0.0f;
    RT_HIP(hipEventElapsedTime(&ms, ev0, ev1));
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);

    sc.nwide = level_base;
    sc.wide.alloc((size_t)sc.nwide * 32u * W);
    RT_HIP(hipMemcpy(sc.wide.p, wide_tmp, (size_t)sc.nwide * 32u * W, hipMemcpyDeviceToDevice));
    sc.wide_depth = depth;
    sc.leaf_size = L;
    sc.stack_bound = (W - 1u) * depth + 1u;
    sc.build_ms = ms;
    if (keep_binary) sc.binary = std::move(kept);
    sc.tris = std::move(dtris);  // world space: the ray queries' refinement reads the vertices
    sc.spheres = std::move(dsph);
}

}  // namespace rtgpu
