"""ctypes binding of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to
check the HIP product path; the product never imports it.  The layouts below
restate the reference's structs independently of the product package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# RT_ORACLE_LIB: another build of rt_oracle.c (the tests' planted-mistake builds use it)
LIB_PATH = Path(os.environ.get("RT_ORACLE_LIB", HERE / "liboracle.so"))

TRIANGLE = np.dtype([("v0", "<f4", 4), ("v1", "<f4", 4), ("v2", "<f4", 4), ("materialIndex", "<u4"),
                     ("modelIndex", "<u4"), ("_pad", "<u4", 2)])
SPHERE = np.dtype([("center", "<f4", 4), ("radius", "<f4"), ("materialIndex", "<u4"), ("modelIndex", "<u4"),
                   ("_pad", "<u4")])
MATERIAL = np.dtype([("albedo", "<f4", 4), ("materialType", "<u4"), ("_pad", "<u4", 3)])
NODE = np.dtype([("minX", "<f4"), ("maxX", "<f4"), ("minY", "<f4"), ("maxY", "<f4"), ("minZ", "<f4"),
                 ("maxZ", "<f4"), ("left", "<u4"), ("right", "<u4"), ("primitiveIndex", "<u4"),
                 ("primitiveType", "<u4")])
PARAMS = np.dtype([("camPos", "<f4", 4), ("camLookAt", "<f4", 4), ("camUpDir", "<f4", 4), ("verticalFOV", "<f4"),
                   ("numTriangles", "<u4"), ("numSpheres", "<u4"), ("numMaterials", "<u4"), ("numLights", "<u4"),
                   ("maxRayTraceDepth", "<u4"), ("randomState", "<u4"), ("_pad", "<u4")])
RAY = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT = np.dtype([("prim", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])
COUNTS = np.dtype([("n_box", "<u4"), ("n_tri", "<u4")])

BRUTE, REF_BVH, WIDE = 0, 1, 2
BUILD_REFERENCE_LEAVES, BUILD_REFERENCE_MORTON = 1, 2
RENDER_REFERENCE_MATERIALS = 1
TRACE_FP32 = 2
NO_HIT = 0xFFFFFFFF
SPHERE_BIT = 0x80000000


def build(force: bool = True) -> Path:
    subprocess.run(["make", "-C", str(HERE)] + (["-B"] if force else []), check=True, capture_output=True)
    return LIB_PATH


if not LIB_PATH.exists():
    build()
lib = C.CDLL(str(LIB_PATH))


class Scene(C.Structure):
    _fields_ = [("tris", C.c_void_p), ("ntri", C.c_uint32), ("sph", C.c_void_p), ("nsph", C.c_uint32),
                ("mats", C.c_void_p), ("nmat", C.c_uint32), ("nodes", C.c_void_p), ("nnodes", C.c_uint32),
                ("slot_prim", C.c_void_p), ("nslots", C.c_uint32), ("wide_nodes", C.c_void_p),
                ("nwide", C.c_uint32), ("width", C.c_uint32), ("wide_root", C.c_uint32)]


_vp, _u32, _f = C.c_void_p, C.c_uint32, C.c_float
for name, res, args in [
    ("orc_rng_seed", _u32, [_u32, _u32, _u32]),
    ("orc_rng_step", _u32, [_u32]),
    ("orc_rng_next", _f, [C.POINTER(_u32)]),
    ("orc_sinf", _f, [_f]),
    ("orc_cosf", _f, [_f]),
    ("orc_f2u_sat", _u32, [_f]),
    ("orc_random_in_unit_sphere", None, [C.POINTER(_u32), _vp]),
    ("orc_random_unit_vector", None, [C.POINTER(_u32), _vp]),
    ("orc_reflect", None, [_vp, _vp, _vp]),
    ("orc_refract", None, [_vp, _vp, _f, _vp]),
    ("orc_schlick", _f, [_f, _f]),
    ("orc_camera", None, [_vp, _u32, _u32, _vp]),
    ("orc_primary_ray", None, [_vp, _u32, _u32, _vp, _vp]),
    ("orc_triangle_hit", C.c_int, [_vp, _vp, _vp, _f, _f, _vp, _vp, _vp, _vp, _vp]),
    ("orc_sphere_hit", C.c_int, [_vp, _vp, _vp, _f, _f, _vp, _vp, _vp]),
    ("orc_aabb_hit_ref", C.c_int, [_vp, _vp, _vp, _vp]),
    ("orc_sphere_uv", None, [_vp, _vp, _vp]),
    ("orc_refine_triangle", C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    ("orc_refine_sphere", C.c_int, [_vp, _vp, _vp, _vp]),
    ("orc_model_to_world", None, [_vp, _u32, _vp, _u32, _vp]),
    ("orc_transform_mat4", None, [_vp, _vp, _vp, _vp]),
    ("orc_enclosing_aabb", None, [_vp, _u32, _vp, _u32, _vp, _vp]),
    ("orc_morton_codes", None, [_vp, _u32, _vp, _u32, _vp, _vp, _u32, _vp]),
    ("orc_stable_radix_sort", None, [_vp, _vp, _u32]),
    ("orc_build_lbvh", C.c_int, [_vp, _u32, _vp, _u32, _u32, _vp, _vp, _vp, _vp, _vp]),
    ("orc_trace_batch", None, [C.POINTER(Scene), C.c_int, _vp, _u32, _u32, _vp, _vp, C.c_int]),
    ("orc_render_rows", None, [C.POINTER(Scene), C.c_int, _vp, _u32, _u32, _u32, _u32, _vp, _u32, _u32,
                               C.POINTER(C.c_uint64), C.c_int]),
    ("orc_resolve", None, [_vp, _u32, _u32, _vp]),
    ("orc_scatter", C.c_int, [_vp, _u32, _vp, _vp, C.c_int, C.POINTER(_u32), _vp, _vp]),
]:
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args


def _p(a):
    return None if a is None else a.ctypes.data


def _f3(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(3))


# ---------------- scalar pieces ----------------
def rng_seed(x, y, rs):
    return lib.orc_rng_seed(x, y, rs)


def rng_sequence(state: int, n: int) -> tuple[np.ndarray, int]:
    st = C.c_uint32(state)
    out = np.array([lib.orc_rng_next(C.byref(st)) for _ in range(n)], dtype=np.float32)
    return out, st.value


def random_in_unit_sphere(state: int):
    st = C.c_uint32(state)
    out = np.zeros(3, np.float32)
    lib.orc_random_in_unit_sphere(C.byref(st), out.ctypes.data)
    return out, st.value


def random_unit_vector(state: int):
    st = C.c_uint32(state)
    out = np.zeros(3, np.float32)
    lib.orc_random_unit_vector(C.byref(st), out.ctypes.data)
    return out, st.value


def camera(params: np.ndarray, w: int, h: int) -> np.ndarray:
    out = np.zeros(12, np.float32)
    p = np.ascontiguousarray(params, dtype=PARAMS)
    lib.orc_camera(p.ctypes.data, w, h, out.ctypes.data)
    return out


def primary_ray(cam: np.ndarray, x: int, y: int):
    o, d = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib.orc_primary_ray(np.ascontiguousarray(cam, np.float32).ctypes.data, x, y, o.ctypes.data, d.ctypes.data)
    return o, d


def triangle_hit(tri: np.ndarray, o, d, tmin, tmax):
    tri = np.ascontiguousarray(np.asarray(tri, dtype=TRIANGLE).reshape(()))
    t, u, v = np.zeros(1, np.float32), np.zeros(1, np.float32), np.zeros(1, np.float32)
    n, bf = np.zeros(3, np.float32), np.zeros(1, np.int32)
    o, d = _f3(o), _f3(d)
    ok = lib.orc_triangle_hit(tri.ctypes.data, o.ctypes.data, d.ctypes.data, tmin, tmax, t.ctypes.data,
                              u.ctypes.data, v.ctypes.data, n.ctypes.data, bf.ctypes.data)
    return bool(ok), float(t[0]), float(u[0]), float(v[0]), n, int(bf[0])


def sphere_hit(sph: np.ndarray, o, d, tmin, tmax):
    sph = np.ascontiguousarray(np.asarray(sph, dtype=SPHERE).reshape(()))
    t, n, bf = np.zeros(1, np.float32), np.zeros(3, np.float32), np.zeros(1, np.int32)
    o, d = _f3(o), _f3(d)
    ok = lib.orc_sphere_hit(sph.ctypes.data, o.ctypes.data, d.ctypes.data, tmin, tmax, t.ctypes.data,
                            n.ctypes.data, bf.ctypes.data)
    return bool(ok), float(t[0]), n, int(bf[0])


def sphere_uv(center):
    c = _f3(center)
    u, v = np.zeros(1, np.float32), np.zeros(1, np.float32)
    lib.orc_sphere_uv(c.ctypes.data, u.ctypes.data, v.ctypes.data)
    return float(u[0]), float(v[0])


def refine_triangle(tri, o, d, t=0.0, u=0.0, v=0.0):
    """The refined (double-precision) t, u, v of a triangle hit; (ok, t, u, v)."""
    tri = np.ascontiguousarray(np.asarray(tri, dtype=TRIANGLE).reshape(()))
    tt, uu, vv = np.float32([t]), np.float32([u]), np.float32([v])
    o, d = _f3(o), _f3(d)
    ok = lib.orc_refine_triangle(tri.ctypes.data, o.ctypes.data, d.ctypes.data, tt.ctypes.data, uu.ctypes.data,
                                 vv.ctypes.data)
    return bool(ok), tt[0], uu[0], vv[0]


def refine_sphere(sph, o, d, t):
    """The refined t of a sphere hit: the double-precision root nearest the fp32 t; (ok, t)."""
    sph = np.ascontiguousarray(np.asarray(sph, dtype=SPHERE).reshape(()))
    tt = np.float32([t])
    o, d = _f3(o), _f3(d)
    ok = lib.orc_refine_sphere(sph.ctypes.data, o.ctypes.data, d.ctypes.data, tt.ctypes.data)
    return bool(ok), tt[0]


def aabb_hit_ref(o, d, bmin, bmax) -> bool:
    a = [_f3(x) for x in (o, d, bmin, bmax)]
    return bool(lib.orc_aabb_hit_ref(*[x.ctypes.data for x in a]))


def transform_mat4(translation, scale, rotation) -> np.ndarray:
    out = np.zeros(16, np.float32)
    tr, sc, rot = _f3(translation), _f3(scale), _f3(rotation)  # keep the arrays alive across the call
    lib.orc_transform_mat4(tr.ctypes.data, sc.ctypes.data, rot.ctypes.data, out.ctypes.data)
    return out


def model_to_world(tris, spheres, models):
    tris = np.ascontiguousarray(tris, dtype=TRIANGLE).copy()
    spheres = np.ascontiguousarray(spheres, dtype=SPHERE).copy()
    models = np.ascontiguousarray(models, dtype=np.float32)
    lib.orc_model_to_world(_p(tris), len(tris), _p(spheres), len(spheres), _p(models))
    return tris, spheres


def enclosing_aabb(tris, spheres):
    tris = np.ascontiguousarray(tris, dtype=TRIANGLE)
    spheres = np.ascontiguousarray(spheres, dtype=SPHERE)
    mn, mx = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib.orc_enclosing_aabb(_p(tris), len(tris), _p(spheres), len(spheres), mn.ctypes.data, mx.ctypes.data)
    return mn, mx


def morton_codes(tris, spheres, emin, emax, flags=0):
    tris = np.ascontiguousarray(tris, dtype=TRIANGLE)
    spheres = np.ascontiguousarray(spheres, dtype=SPHERE)
    out = np.zeros(len(tris) + len(spheres), np.uint32)
    mn, mx = _f3(emin), _f3(emax)
    lib.orc_morton_codes(_p(tris), len(tris), _p(spheres), len(spheres), mn.ctypes.data, mx.ctypes.data, flags,
                         out.ctypes.data)
    return out


def stable_radix_sort(keys, vals):
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib.orc_stable_radix_sort(k.ctypes.data, v.ctypes.data, len(k))
    return k, v


def build_lbvh(tris, spheres=None, flags=0):
    """Reference builder output: (nodes[2N-1], parent[2N-1], slot_prim[N], sorted_codes, sorted_prims)."""
    tris = np.ascontiguousarray(tris, dtype=TRIANGLE)
    spheres = np.ascontiguousarray(spheres if spheres is not None else np.zeros(0, SPHERE), dtype=SPHERE)
    n = len(tris) + len(spheres)
    nodes = np.zeros(max(2 * n - 1, 0), NODE)
    parent = np.zeros(max(2 * n - 1, 0), np.uint32)
    slots = np.zeros(n, np.uint32)
    codes = np.zeros(n, np.uint32)
    prims = np.zeros(n, np.uint32)
    lib.orc_build_lbvh(_p(tris), len(tris), _p(spheres), len(spheres), flags, _p(nodes), _p(parent), _p(slots),
                       _p(codes), _p(prims))
    return nodes, parent, slots, codes, prims


class OracleScene:
    """Keeps the numpy arrays alive behind an orc_scene."""

    def __init__(self, tris, spheres=None, materials=None, nodes=None, slot_prim=None, wide=None, width=4):
        self.tris = np.ascontiguousarray(tris, dtype=TRIANGLE)
        self.spheres = np.ascontiguousarray(spheres if spheres is not None else np.zeros(0, SPHERE), dtype=SPHERE)
        self.mats = np.ascontiguousarray(materials if materials is not None else np.zeros(0, MATERIAL),
                                         dtype=MATERIAL)
        if slot_prim is None:
            nodes, _, slot_prim, _, _ = build_lbvh(self.tris, self.spheres)
        self.nodes = np.ascontiguousarray(nodes if nodes is not None else np.zeros(0, NODE), dtype=NODE)
        self.slots = np.ascontiguousarray(slot_prim, dtype=np.uint32)
        self.wide = None if wide is None else np.ascontiguousarray(wide, dtype=np.uint32)
        self.width = width
        self.c = Scene(_p(self.tris), len(self.tris), _p(self.spheres), len(self.spheres), _p(self.mats),
                       len(self.mats), _p(self.nodes), len(self.nodes), _p(self.slots), len(self.slots),
                       _p(self.wide), 0 if self.wide is None else len(self.wide) // (8 * width), width, 0)

    def trace(self, rays, method=BRUTE, any_hit=False, counts=False, threads=0, fp32=False):
        """Ray queries.  t, u, v are refined in double precision (the product's default ray-query
        output) unless fp32 is set, which returns the reference's fp32 triangleHit values."""
        rays = np.ascontiguousarray(rays, dtype=RAY)
        hits = np.zeros(len(rays), HIT)
        cnt = np.zeros(len(rays), COUNTS) if counts else None
        flags = (1 if any_hit else 0) | (TRACE_FP32 if fp32 else 0)
        lib.orc_trace_batch(C.byref(self.c), method, _p(rays), len(rays), flags, _p(hits), _p(cnt), threads)
        return (hits, cnt) if counts else hits

    def render(self, params, width, height, spp, accum=None, flags=0, method=BRUTE, rows=None, threads=0):
        params = np.ascontiguousarray(params, dtype=PARAMS)
        if accum is None:
            accum = np.zeros((height * width, 4), np.float32)
            accum[:, 3] = 1.0
        accum = np.ascontiguousarray(accum, dtype=np.float32)
        r0, r1 = rows if rows is not None else (0, height)
        segs = C.c_uint64(0)
        lib.orc_render_rows(C.byref(self.c), method, params.ctypes.data, width, height, spp, flags, _p(accum), r0,
                            r1, C.byref(segs), threads)
        return accum, segs.value


def scatter(material, d, n, backface: bool, state: int, flags: int = 0):
    """One bounce's scatter: (continues, attenuation, out direction, state after its draws)."""
    m = np.ascontiguousarray(np.asarray(material, dtype=MATERIAL).reshape(()))
    d, n = _f3(d), _f3(n)
    st = C.c_uint32(state)
    att, out = np.zeros(3, np.float32), np.zeros(3, np.float32)
    ok = lib.orc_scatter(m.ctypes.data, flags, d.ctypes.data, n.ctypes.data, 1 if backface else 0, C.byref(st),
                         att.ctypes.data, out.ctypes.data)
    return bool(ok), att, out, st.value


def resolve(accum, spp):
    accum = np.ascontiguousarray(accum, dtype=np.float32)
    out = np.zeros_like(accum)
    lib.orc_resolve(_p(accum), len(accum), spp, _p(out))
    return out
