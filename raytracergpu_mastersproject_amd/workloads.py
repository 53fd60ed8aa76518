"""Synthetic workloads (scenes, ray batches, render setups) with explicit seeds.

Data PRNG: SplitMix64 (Steele, Lea, Flood 2014), counter based, so every array
is reproducible on any machine.  Geometry is generated in float64 and rounded
to float32 once.

Workloads:
  soup_scene       random triangle soup + ground quad + emissive quad (1,024 triangles, 256x256, 1 spp)
  icosphere_scene  noise-displaced geodesic icosphere, per-triangle materials
                   (60% lambertian / 30% metal / 10% dielectric) + emissive quad
  cluster_scene    triangle soup in Gaussian clusters (64 x 65,536 = 4,194,304 triangles)
  incoherent_rays  origins uniform in the scene box, directions uniform on the sphere
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from . import layouts as L

_GEN_PATH = Path(__file__).resolve().parent / "libworkloads.so"
_gen = None


def _genlib():
    """The C++ generators (host/workloads_gen.cpp) used for the large workloads."""
    global _gen
    if _gen is None:
        if not _GEN_PATH.exists():
            raise ImportError(f"{_GEN_PATH} not built: run `python -m raytracergpu_mastersproject_amd.build`")
        g = C.CDLL(str(_GEN_PATH))
        vp, u32, u64, dbl, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_double, C.c_int
        g.rtw_uniform.argtypes = [u64, u64, u64, dbl, dbl, vp, i32]
        g.rtw_u64.argtypes = [u64, u64, u64, vp, i32]
        g.rtw_cluster_soup.argtypes = [u32, u32, dbl, dbl, u64, dbl, dbl, u32, vp, i32]
        g.rtw_incoherent_rays.argtypes = [u64, vp, vp, u64, C.c_float, C.c_float, vp, i32]
        g.rtw_bounds.argtypes = [vp, u64, vp, vp]
        for f in (g.rtw_uniform, g.rtw_u64, g.rtw_cluster_soup, g.rtw_incoherent_rays, g.rtw_bounds):
            f.restype = None
        _gen = g
    return _gen


def _threads():
    return int(os.environ.get("OMP_NUM_THREADS", 0) or 0)

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-based SplitMix64: value i of a stream is mix(seed + (i+1)*gamma)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed)
        self.count = 0

    def u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            i = np.arange(self.count + 1, self.count + 1 + n, dtype=np.uint64)
            z = self.state + i * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.count += n
        return z

    def uniform(self, n: int, lo=0.0, hi=1.0) -> np.ndarray:
        """float64 uniform in [lo, hi) from the top 53 bits."""
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int) -> np.ndarray:
        """Standard normal by Box-Muller (two uniforms per value)."""
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)

    def sphere_dirs(self, n: int) -> np.ndarray:
        """Directions uniform on S^2: z ~ U[-1,1], phi ~ U[0, 2pi)."""
        z = self.uniform(n, -1.0, 1.0)
        phi = self.uniform(n, 0.0, 2.0 * np.pi)
        r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def make_triangles(v0, v1, v2, material=0) -> np.ndarray:
    n = len(v0)
    t = np.zeros(n, dtype=L.TRIANGLE)
    t["v0"][:, :3] = v0
    t["v1"][:, :3] = v1
    t["v2"][:, :3] = v2
    t["v0"][:, 3] = t["v1"][:, 3] = t["v2"][:, 3] = 1.0
    t["materialIndex"] = material
    return t


def quad(y, x0, x1, z0, z1, material, flip=False) -> np.ndarray:
    a, b, c, d = (x0, y, z0), (x1, y, z0), (x1, y, z1), (x0, y, z1)
    if flip:
        return make_triangles(np.array([a, a]), np.array([c, d]), np.array([b, c]), material)
    return make_triangles(np.array([a, a]), np.array([b, c]), np.array([c, d]), material)


def make_materials(albedo, mtype, extra=None) -> np.ndarray:
    albedo = np.asarray(albedo, dtype=np.float64).reshape(-1, 3)
    m = np.zeros(len(albedo), dtype=L.MATERIAL)
    m["albedo"][:, :3] = albedo
    m["materialType"] = mtype
    if extra is not None:
        m["albedo"][:, 3] = extra
    return m


def random_soup(rng: SplitMix64, n: int, lo=-1.0, hi=1.0, radius=0.0375) -> tuple:
    c = rng.uniform(3 * n, lo, hi).reshape(n, 3)
    d = rng.sphere_dirs(3 * n).reshape(n, 3, 3)
    return c + radius * d[:, 0], c + radius * d[:, 1], c + radius * d[:, 2]


class RenderSetup(dict):
    """tris, spheres, materials, params (ParameterUBO), width, height, spp."""


def soup_scene(ntri: int = 1024, seed: int = 0, width: int = 256, height: int = 256, spp: int = 1,
               depth: int = 8, random_state: int = 0) -> RenderSetup:
    """1,024 random fp32 triangles (centroids U[-1,1]^3, mean edge ~0.05) + ground quad at y=-1.1
    + an emissive quad at y=3 (the BVH kernel's background is 0, raytraceBVH.comp:52), all lambertian
    (0.73), camera (0,0.5,-3.5) -> origin, vfov 40."""
    rng = SplitMix64(seed)
    v0, v1, v2 = random_soup(rng, ntri)
    soup = make_triangles(v0, v1, v2, 0)
    ground = quad(-1.1, -3.0, 3.0, -3.0, 3.0, 0)
    light = quad(3.0, -1.5, 1.5, -1.5, 1.5, 1, flip=True)
    tris = np.concatenate([soup, ground, light])
    mats = np.concatenate([make_materials([0.73, 0.73, 0.73], L.DIFFUSE), make_materials([4.0, 4.0, 4.0], L.LIGHT)])
    params = L.make_params((0.0, 0.5, -3.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, len(tris), 0, len(mats),
                           depth, random_state)
    return RenderSetup(tris=tris, spheres=np.zeros(0, L.SPHERE), materials=mats, params=params, width=width,
                       height=height, spp=spp)


def _hash3(ix, iy, iz, seed):
    with np.errstate(over="ignore"):
        h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841)
             ^ iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^ np.uint64(seed) * _GAMMA)
        h = (h ^ (h >> np.uint64(30))) * _M1
        h = (h ^ (h >> np.uint64(27))) * _M2
        h = h ^ (h >> np.uint64(31))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def value_noise(p: np.ndarray, freq=4.0, seed=1) -> np.ndarray:
    """Trilinear value noise in [-1, 1] on a hashed integer lattice, smoothstep-weighted."""
    q = p * freq
    i = np.floor(q).astype(np.int64)
    f = q - i
    w = f * f * (3.0 - 2.0 * f)
    out = np.zeros(len(p))
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                h = _hash3(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
                wx = w[:, 0] if dx else 1.0 - w[:, 0]
                wy = w[:, 1] if dy else 1.0 - w[:, 1]
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out += wx * wy * wz * h
    return 2.0 * out - 1.0


def icosphere(freq: int) -> tuple:
    """Geodesic icosphere: each icosahedron face split into freq^2 triangles, projected to r=1."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)])
    f = freq
    ii, jj = np.meshgrid(np.arange(f), np.arange(f), indexing="ij")
    up = ii + jj <= f - 1
    dn = ii + jj <= f - 2
    iu, ju = ii[up], jj[up]
    idn, jdn = ii[dn], jj[dn]

    def pt(A, B, Cc, i, j):
        return A + (B - A) * (i[:, None] / f) + (Cc - A) * (j[:, None] / f)

    out0, out1, out2 = [], [], []
    for a, b, c in F:
        A, B, Cc = V[a], V[b], V[c]
        out0.append(pt(A, B, Cc, iu, ju)); out1.append(pt(A, B, Cc, iu + 1, ju)); out2.append(pt(A, B, Cc, iu, ju + 1))
        out0.append(pt(A, B, Cc, idn + 1, jdn)); out1.append(pt(A, B, Cc, idn + 1, jdn + 1))
        out2.append(pt(A, B, Cc, idn, jdn + 1))
    P = [np.concatenate(o) for o in (out0, out1, out2)]
    return tuple(p / np.linalg.norm(p, axis=1, keepdims=True) for p in P)


def icosphere_scene(freq: int = 224, seed: int = 1, width: int = 1920, height: int = 1080, spp: int = 16,
                    depth: int = 8, random_state: int = 1) -> RenderSetup:
    """Noise-displaced icosphere (20*freq^2 triangles; freq 224 -> 1,003,520), r = 1 + 0.05*noise(p),
    per-triangle material by hash: 60% lambertian (albedo U[0.2,0.9]^3), 30% metal (albedo U[0.5,1]^3,
    fuzz U[0,0.3]), 10% dielectric (ior 1.5); plus one emissive quad above.  Camera (0,0.6,-3.4) ->
    origin, vfov 40."""
    P0, P1, P2 = icosphere(freq)
    disp = [p * (1.0 + 0.05 * value_noise(p, 4.0, seed))[:, None] for p in (P0, P1, P2)]
    n = len(P0)
    rng = SplitMix64(seed)
    kind = rng.uniform(n)
    col = rng.uniform(3 * n).reshape(n, 3)
    fuzz = rng.uniform(n, 0.0, 0.3)
    mtype = np.where(kind < 0.6, L.DIFFUSE, np.where(kind < 0.9, L.METALLIC, L.DIELECTRIC)).astype(np.uint32)
    albedo = np.where((mtype == L.DIFFUSE)[:, None], 0.2 + 0.7 * col, 0.5 + 0.5 * col)
    albedo[mtype == L.DIELECTRIC] = 1.0
    extra = np.where(mtype == L.METALLIC, fuzz, np.where(mtype == L.DIELECTRIC, 1.5, 0.0))
    mats = np.concatenate([make_materials(albedo, 0, extra), make_materials([4.0, 4.0, 4.0], L.LIGHT)])
    mats["materialType"][:n] = mtype
    tris = make_triangles(*disp, np.arange(n, dtype=np.uint32))
    light = quad(2.5, -2.0, 2.0, -2.0, 2.0, n, flip=True)
    tris = np.concatenate([tris, light])
    params = L.make_params((0.0, 0.6, -3.4), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, len(tris), 0, len(mats),
                           depth, random_state)
    return RenderSetup(tris=tris, spheres=np.zeros(0, L.SPHERE), materials=mats, params=params, width=width,
                       height=height, spp=spp)


def cluster_scene(nclusters: int = 64, per_cluster: int = 65536, sigma: float = 0.05, tri_radius: float = 0.005,
                  seed: int = 2, centres_lo=-1.0, centres_hi=1.0) -> tuple:
    """Triangle soup in Gaussian clusters: centres ~ U[-1,1]^3, centroids ~ N(centre, sigma^2 I),
    vertices centroid + tri_radius * (uniform unit vectors).  Returns (tris, materials)."""
    n = nclusters * per_cluster
    tris = np.zeros(n, dtype=L.TRIANGLE)
    _genlib().rtw_cluster_soup(nclusters, per_cluster, sigma, tri_radius, seed, centres_lo, centres_hi, 0,
                               tris.ctypes.data, _threads())
    return tris, make_materials([0.73, 0.73, 0.73], L.DIFFUSE)


def scene_bounds(tris: np.ndarray) -> tuple:
    lo, hi = np.zeros(3), np.zeros(3)
    tris = np.ascontiguousarray(tris, dtype=L.TRIANGLE)
    _genlib().rtw_bounds(tris.ctypes.data, len(tris), lo.ctypes.data, hi.ctypes.data)
    return lo, hi


def incoherent_rays(n: int, lo, hi, seed: int = 3, tmin=0.001, tmax=1e7) -> np.ndarray:
    """Origins uniform in the box [lo, hi], directions uniform on the sphere."""
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    r = np.zeros(n, dtype=L.RAY)
    _genlib().rtw_incoherent_rays(n, lo.ctypes.data, hi.ctypes.data, seed, tmin, tmax, r.ctypes.data, _threads())
    return r


def ray_query_workload(nrays: int = 16_777_216, nclusters: int = 64, per_cluster: int = 65536) -> tuple:
    """16M incoherent rays vs 4M-triangle soup in 64 Gaussian clusters (sigma 0.05)."""
    tris, mats = cluster_scene(nclusters, per_cluster, seed=2)
    lo, hi = scene_bounds(tris)
    rays = incoherent_rays(nrays, lo, hi, seed=3)
    return tris, mats, rays


def batched_workload(nscenes=256, tri_lo=1_000, tri_hi=2_000_000, ray_lo=4096, ray_hi=1_048_576,
                     clusters=8, seed_tris=4, seed_rays=5, scale=1.0):
    """Many scenes with log-uniform triangle counts (8 Gaussian clusters each) and log-uniform ray
    batches.  `scale` shrinks all counts (tests)."""
    rt = SplitMix64(seed_tris)
    rr = SplitMix64(seed_rays)
    ntris = np.exp(rt.uniform(nscenes, np.log(tri_lo), np.log(tri_hi))) * scale
    nrays = np.exp(rr.uniform(nscenes, np.log(ray_lo), np.log(ray_hi))) * scale
    out = []
    for k in range(nscenes):
        nt = max(clusters, int(ntris[k]) // clusters * clusters)
        tris, mats = cluster_scene(clusters, nt // clusters, seed=1000 + k)
        lo, hi = scene_bounds(tris)
        rays = incoherent_rays(max(1, int(nrays[k])), lo, hi, seed=100000 + k)
        out.append((tris, mats, rays))
    return out


def material_palette(seed: int = 6) -> np.ndarray:
    """40 materials: 24 lambertian (albedo U[0.2,0.9]^3), 12 metal (albedo U[0.5,1]^3, fuzz U[0,0.3]),
    4 dielectric (ior 1.5), then one emitter (albedo 4) at index 40.  A triangle drawing its index
    uniformly from 0..39 is 60% lambertian, 30% metal, 10% dielectric."""
    rng = SplitMix64(seed)
    col = rng.uniform(3 * 40).reshape(40, 3)
    fuzz = rng.uniform(40, 0.0, 0.3)
    m = np.concatenate([make_materials(0.2 + 0.7 * col[:24], L.DIFFUSE),
                        make_materials(0.5 + 0.5 * col[24:36], L.METALLIC, fuzz[24:36]),
                        make_materials(np.ones((4, 3)), L.DIELECTRIC, np.full(4, 1.5)),
                        make_materials([4.0, 4.0, 4.0], L.LIGHT)])
    return m


def served_scene(ntri: int, seed: int, nclusters: int = 8) -> RenderSetup:
    """A clustered soup of about `ntri` triangles (8 Gaussian clusters, sigma 0.05, centres U[-1,1]^3)
    with palette materials, a lambertian ground quad (y = -1.1), an emissive quad (y = 3), and the
    soup scene's camera ((0,0.5,-3.5) -> origin, vfov 40)."""
    per = max(1, ntri // nclusters)
    tris, _ = cluster_scene(nclusters, per, seed=seed)
    tris["materialIndex"] = (SplitMix64(seed ^ 0x5EED).u64(len(tris)) % np.uint64(40)).astype(np.uint32)
    mats = material_palette()
    ground = quad(-1.1, -3.0, 3.0, -3.0, 3.0, 0)
    light = quad(3.0, -1.5, 1.5, -1.5, 1.5, 40, flip=True)
    tris = np.concatenate([tris, ground, light])
    params = L.make_params((0.0, 0.5, -3.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, len(tris), 0, len(mats), 8, 0)
    return RenderSetup(tris=tris, spheres=np.zeros(0, L.SPHERE), materials=mats, params=params, width=0, height=0,
                       spp=0)


def serving_scene_sizes(nscenes: int = 64, lo: float = 1e4, hi: float = 4_194_304, seed: int = 6) -> np.ndarray:
    """Triangle counts log-uniform in [lo, hi]."""
    return np.exp(SplitMix64(seed).uniform(nscenes, np.log(lo), np.log(hi))).astype(np.int64)


SERVING_RESOLUTIONS = ((256, 256), (512, 512), (1920, 1080))


def serving_requests(n: int, nscenes: int = 64, max_spp: int = 64, seed: int = 7, scale: float = 1.0):
    """Request descriptors (scene, width, height, spp, seed = request index): scene U{0..nscenes-1},
    resolution U{256^2, 512^2, 1920x1080} (scaled by `scale` for tests), spp U{1..max_spp}, depth 8."""
    rng = SplitMix64(seed)
    u = rng.u64(3 * n).reshape(n, 3)
    out = []
    for i in range(n):
        w, h = SERVING_RESOLUTIONS[int(u[i, 1] % np.uint64(3))]
        out.append({"scene": int(u[i, 0] % np.uint64(nscenes)), "width": max(8, int(w * scale)),
                    "height": max(8, int(h * scale)), "spp": 1 + int(u[i, 2] % np.uint64(max_spp)), "seed": i})
    return out


def poisson_arrivals(n: int, rate_per_s: float, seed: int = 8) -> np.ndarray:
    """Arrival times (ms) of a Poisson process of the given rate."""
    gaps = -np.log1p(-SplitMix64(seed).uniform(n)) / rate_per_s
    return np.cumsum(gaps) * 1000.0
