"""Pins for the oracle's refined ray-query results (orc_refine_triangle / orc_refine_sphere / orc_sphere_uv).

The ray query decides which primitive a ray hits with the reference's fp32 triangleHit
(raytraceBVH.comp:118-149) and then reports t, u, v recomputed in double precision.  The pins:
- triangles: exact rational arithmetic (Cramer's rule on the fp32 inputs with Python fractions)
  gives the true intersection; the refined values must be within one fp32 ulp of it.  At the
  ray-query workload's scale (0.005-unit triangles a unit or two away) the fp32 triangleHit values
  themselves miss the 1e-5 bar on a large share of hits, so the sample is one where the check bites;
- spheres: roots of the quadratic in 60-digit decimal arithmetic, and the choice of root (the one
  nearest the fp32 t: the near root from outside, the far root from inside);
- sphere u, v (raytraceBVH.comp:172-173, from the centre): known angles and the acos domain.
"""
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np

import oracle as O


def _tri(v0, v1, v2):
    t = np.zeros((), O.TRIANGLE)
    t["v0"][:3], t["v1"][:3], t["v2"][:3] = v0, v1, v2
    return t


def _det3(a, b, c):
    return (a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0])
            + a[2] * (b[0] * c[1] - b[1] * c[0]))


def exact_intersection(v0, v1, v2, o, d):
    """(t, u, v) with o + t d = v0 + u (v1 - v0) + v (v2 - v0), exactly, by Cramer's rule."""
    F = [[Fraction(float(x)) for x in p] for p in (v0, v1, v2, o, d)]
    v0, v1, v2, o, d = F
    e1 = [v1[k] - v0[k] for k in range(3)]
    e2 = [v2[k] - v0[k] for k in range(3)]
    nd = [-x for x in d]
    rhs = [o[k] - v0[k] for k in range(3)]
    den = _det3(nd, e1, e2)
    return _det3(rhs, e1, e2) / den, _det3(nd, rhs, e2) / den, _det3(nd, e1, rhs) / den


def _headline_scale_hits(n, seed):
    """fp32 hits of 0.005-unit triangles (the ray-query workload's size) by rays from up to two units away."""
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        c = rng.uniform(-1, 1, 3)
        dirs = rng.normal(size=(3, 3))
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
        v = (c + 0.005 * dirs).astype(np.float32)
        o = rng.uniform(-1, 1, 3).astype(np.float32)
        a, b = rng.uniform(0.02, 0.96, 2)
        if a + b > 0.98:
            continue
        target = v[0] + a * (v[1] - v[0]) + b * (v[2] - v[0])
        d = (target - o)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        ok, t, u, w, _, _ = O.triangle_hit(_tri(*v), o, d, 0.001, 1e7)
        if ok:
            out.append((v, o, d, np.float32(t), np.float32(u), np.float32(w)))
    return out


def _ulp(x):
    return float(np.spacing(np.float32(abs(x))))


def test_refined_triangle_hits_are_the_exact_intersection_rounded():
    raw_over = 0
    hits = _headline_scale_hits(600, seed=1)
    for v, o, d, t32, u32, w32 in hits:
        ok, t, u, w = O.refine_triangle(_tri(*v), o, d, t32, u32, w32)
        assert ok
        et, eu, ew = exact_intersection(v[0], v[1], v[2], o, d)
        for got, ex in ((t, et), (u, eu), (w, ew)):
            assert abs(float(got) - float(ex)) <= max(_ulp(got), 1e-12), (got, float(ex))
        assert abs(float(t) - float(et)) <= 1e-5 * abs(float(et))
        assert abs(float(u) - float(eu)) <= 1e-5 and abs(float(w) - float(ew)) <= 1e-5
        raw_over += (abs(float(t32) - float(et)) > 1e-5 * abs(float(et)) or abs(float(u32) - float(eu)) > 1e-5
                     or abs(float(w32) - float(ew)) > 1e-5)
    # the fp32 plane-based values miss the bar on a large share of these hits: the sample bites
    assert raw_over > 0.05 * len(hits)


def test_refined_triangle_on_ill_conditioned_and_degenerate_cases():
    # grazing ray (n.d just above triangleHit's 1e-4 cutoff) on a unit triangle
    v = np.float32([[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    d = np.float32([0.0, 0.99999, -0.0045])
    d = (d / np.linalg.norm(d)).astype(np.float32)
    o = (np.float32([0.2, 0.3, 0.0]) - d).astype(np.float32)  # one unit before a point inside
    ok32, t32, u32, w32, _, _ = O.triangle_hit(_tri(*v), o, d, 0.001, 1e7)
    assert ok32
    ok, t, u, w = O.refine_triangle(_tri(*v), o, d, t32, u32, w32)
    et, eu, ew = exact_intersection(v[0], v[1], v[2], o, d)
    assert ok and abs(float(t) - float(et)) <= _ulp(t) and abs(float(w) - float(ew)) <= _ulp(w)
    # a zero-area triangle has no solution: the inputs are returned untouched
    z = np.float32([[0, 0, 0], [1, 1, 1], [2, 2, 2]])
    ok, t, u, w = O.refine_triangle(_tri(*z), o, d, 5.0, 0.25, 0.5)
    assert not ok and (t, u, w) == (5.0, 0.25, 0.5)


def _sphere(c, r):
    s = np.zeros((), O.SPHERE)
    s["center"][:3] = c
    s["center"][3] = 1.0
    s["radius"] = r
    return s


def _decimal_roots(c, r, o, d):
    getcontext().prec = 60
    C, O_, D = ([Decimal(float(x)) for x in v] for v in (c, o, d))
    R = Decimal(float(r))
    oc = [O_[k] - C[k] for k in range(3)]
    a = sum(D[k] * D[k] for k in range(3))
    hb = sum(oc[k] * D[k] for k in range(3))
    cc = sum(oc[k] * oc[k] for k in range(3)) - R * R
    sq = (hb * hb - a * cc).sqrt()
    return float((-hb - sq) / a), float((-hb + sq) / a)


def test_refined_sphere_roots():
    rng = np.random.default_rng(3)
    for k in range(300):
        c = rng.uniform(-1, 1, 3).astype(np.float32)
        r = np.float32(rng.uniform(0.001, 0.5))
        u3 = rng.normal(size=(2, 3))
        u3 /= np.linalg.norm(u3, axis=1, keepdims=True)
        inside = k % 3 == 0
        o = (c + (0.5 * r if inside else 3.0) * u3[0]).astype(np.float32)
        target = c + 0.7 * r * u3[1]
        d = (target - o)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        ok32, t32, _, _ = O.sphere_hit(_sphere(c, r), o, d, 0.0, 1e7)
        if not ok32:
            continue
        ok, t = O.refine_sphere(_sphere(c, r), o, d, t32)
        near, far = _decimal_roots(c, r, o, d)
        want = far if inside else near  # from inside only the far root is ahead of the ray
        assert ok and abs(float(t) - want) <= _ulp(want), (k, t, want, t32)
    # no real root in double: t is returned untouched
    ok, t = O.refine_sphere(_sphere([0, 0, 0], 1.0), [0, 2, 0], [1, 0, 0], 7.0)
    assert not ok and t == np.float32(7.0)


def test_sphere_uv_known_angles():
    # u = (atan(-c.z, c.x) + pi) / (2 pi), v = acos(-c.y) / pi
    assert O.sphere_uv([1, 0, 0]) == (0.5, 0.5)
    assert O.sphere_uv([0, 0, -1]) == (0.75, 0.5)
    assert O.sphere_uv([0, 0, 1]) == (0.25, 0.5)
    assert O.sphere_uv([-1, 0, 0])[0] == 0.0       # atan2(-0, -1) = -pi
    assert O.sphere_uv([0, -1, 0])[1] == 0.0       # acos(1) = 0
    assert O.sphere_uv([0, 1, 0])[1] == 1.0        # acos(-1) = pi
    u, v = O.sphere_uv([0.5, 0.5, -0.5])
    assert u == np.float32(0.625) and v == np.float32(2.0 / 3.0)  # atan2(0.5, 0.5) = pi/4, acos(-0.5) = 2pi/3
    assert np.isnan(O.sphere_uv([0, -1000, 0])[1])  # the reference's ground sphere: acos outside [-1, 1]


def test_traced_hits_are_refined_and_keep_the_fp32_ids():
    """OracleScene.trace refines by default; the ids are the fp32 traversal's."""
    rng = np.random.default_rng(4)
    n = 3000
    c = rng.uniform(-0.08, 0.08, (n, 3))
    dirs = rng.normal(size=(n, 3, 3))
    dirs /= np.linalg.norm(dirs, axis=2, keepdims=True)
    tris = np.zeros(n, O.TRIANGLE)
    for k, name in enumerate(("v0", "v1", "v2")):
        tris[name][:, :3] = c + 0.005 * dirs[:, k]
    mats = np.zeros(1, O.MATERIAL)
    sc = O.OracleScene(tris, None, mats)
    rays = np.zeros(4000, O.RAY)
    rays["o"] = rng.uniform(-0.1, 0.1, (len(rays), 3))
    dd = rng.normal(size=(len(rays), 3))
    rays["d"] = dd / np.linalg.norm(dd, axis=1, keepdims=True)
    rays["tmin"], rays["tmax"] = 0.001, 1e7
    raw = sc.trace(rays, O.REF_BVH, fp32=True)
    ref = sc.trace(rays, O.REF_BVH)
    assert np.array_equal(raw["prim"], ref["prim"])
    idx = np.nonzero(ref["prim"] != O.NO_HIT)[0]
    assert len(idx) > 100
    for i in idx[:200]:
        t = tris[ref["prim"][i]]
        et, eu, ew = exact_intersection(t["v0"][:3], t["v1"][:3], t["v2"][:3], rays["o"][i], rays["d"][i])
        for got, ex in ((ref["t"][i], et), (ref["u"][i], eu), (ref["v"][i], ew)):
            assert abs(float(got) - float(ex)) <= max(_ulp(got), 1e-12)
