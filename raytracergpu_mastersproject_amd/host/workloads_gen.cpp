// workloads_gen.cpp -- multithreaded generators for the large synthetic workloads.
//
// SplitMix64 (Steele, Lea, Flood 2014) used counter-style: value i (0-based) of the stream
// with seed s is mix(s + (i+1) * 0x9E3779B97F4A7C15), identical to workloads.SplitMix64.
// Uniform doubles take the top 53 bits.  Every output element depends only on its index, so
// chunks are generated in parallel with identical results for any thread count.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "rtgpu.h"

namespace {

constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ull;

inline uint64_t mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
inline uint64_t sm_u64(uint64_t seed, uint64_t i) { return mix(seed + (i + 1) * kGamma); }
inline double sm_uniform(uint64_t seed, uint64_t i) {
    return (double)(sm_u64(seed, i) >> 11) * (1.0 / 9007199254740992.0);
}

template <class F> void parallel_for(uint64_t n, int threads, F &&f) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = (int)std::min<uint64_t>((uint64_t)threads, std::max<uint64_t>(1, n / 4096));
    if (threads <= 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> ts;
    uint64_t chunk = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        uint64_t b = t * chunk, e = std::min(n, b + chunk);
        if (b < e) ts.emplace_back([&f, b, e] { f(b, e); });
    }
    for (auto &t : ts) t.join();
}

// z ~ U[-1,1], phi ~ U[0, 2pi): direction uniform on the sphere (draws k0 and k0+1 of `seed` at offset)
inline void sphere_dir(double z, double phi, double out[3]) {
    double r = std::sqrt(std::max(0.0, 1.0 - z * z));
    out[0] = r * std::cos(phi);
    out[1] = r * std::sin(phi);
    out[2] = z;
}

}  // namespace

extern "C" {

// n uniform doubles in [lo, hi): values offset .. offset+n-1 of the stream
void rtw_uniform(uint64_t seed, uint64_t offset, uint64_t n, double lo, double hi, double *out, int threads) {
    parallel_for(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) out[i] = lo + (hi - lo) * sm_uniform(seed, offset + i);
    });
}

void rtw_u64(uint64_t seed, uint64_t offset, uint64_t n, uint64_t *out, int threads) {
    parallel_for(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) out[i] = sm_u64(seed, offset + i);
    });
}

// Clustered triangle soup: centres ~ U[lo,hi]^3 (draws 0..3C-1), then per triangle t of cluster
// c = t / per_cluster: centroid = centre + sigma * N(0,1)^3 (Box-Muller, draws from stream seed+1),
// vertices = centroid + tri_radius * (3 directions uniform on the sphere, stream seed+2).
void rtw_cluster_soup(uint32_t nclusters, uint32_t per_cluster, double sigma, double tri_radius, uint64_t seed,
                      double lo, double hi, uint32_t material, rt_triangle *out, int threads) {
    std::vector<double> centres(3ull * nclusters);
    for (uint64_t i = 0; i < 3ull * nclusters; ++i) centres[i] = lo + (hi - lo) * sm_uniform(seed, i);
    const uint64_t n = (uint64_t)nclusters * per_cluster;
    const uint64_t sg = seed + 1, sd = seed + 2;
    parallel_for(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t t = b; t < e; ++t) {
            const double *c = &centres[3 * (t / per_cluster)];
            double cen[3];
            for (int a = 0; a < 3; ++a) {
                double u1 = sm_uniform(sg, 6 * t + 2 * a), u2 = sm_uniform(sg, 6 * t + 2 * a + 1);
                double g = std::sqrt(-2.0 * std::log1p(-u1)) * std::cos(2.0 * M_PI * u2);
                cen[a] = c[a] + sigma * g;
            }
            rt_triangle tri;
            std::memset(&tri, 0, sizeof tri);
            float *vs[3] = {tri.v0, tri.v1, tri.v2};
            for (int k = 0; k < 3; ++k) {
                double z = -1.0 + 2.0 * sm_uniform(sd, 6 * t + 2 * k);
                double phi = 2.0 * M_PI * sm_uniform(sd, 6 * t + 2 * k + 1);
                double d[3];
                sphere_dir(z, phi, d);
                for (int a = 0; a < 3; ++a) vs[k][a] = (float)(cen[a] + tri_radius * d[a]);
                vs[k][3] = 1.0f;
            }
            tri.materialIndex = material;
            out[t] = tri;
        }
    });
}

// Incoherent rays: origin ~ U(box) (stream seed, draws 3i..3i+2), direction uniform on the
// sphere (stream seed+1, draws 2i, 2i+1).
void rtw_incoherent_rays(uint64_t n, const double *lo, const double *hi, uint64_t seed, float tmin, float tmax,
                         rt_ray *out, int threads) {
    parallel_for(n, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            rt_ray r;
            for (int a = 0; a < 3; ++a) r.o[a] = (float)(lo[a] + (hi[a] - lo[a]) * sm_uniform(seed, 3 * i + a));
            double z = -1.0 + 2.0 * sm_uniform(seed + 1, 2 * i), phi = 2.0 * M_PI * sm_uniform(seed + 1, 2 * i + 1);
            double d[3];
            sphere_dir(z, phi, d);
            for (int a = 0; a < 3; ++a) r.d[a] = (float)d[a];
            r.tmin = tmin;
            r.tmax = tmax;
            out[i] = r;
        }
    });
}

// bounding box of all triangle vertices
void rtw_bounds(const rt_triangle *tris, uint64_t n, double *lo, double *hi) {
    for (int a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; }
    for (uint64_t i = 0; i < n; ++i) {
        const float *vs[3] = {tris[i].v0, tris[i].v1, tris[i].v2};
        for (const float *v : vs)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], (double)v[a]);
                hi[a] = std::max(hi[a], (double)v[a]);
            }
    }
}

}  // extern "C"
