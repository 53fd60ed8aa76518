"""Pins for the oracle's intersection, camera and transform functions
(raytraceBVH.comp:49-81, 118-193, 329-342; TransformComponent.cpp:9-45; ModelSpaceToWorldSpace.comp:32-48)."""
import math

import numpy as np

import oracle as O


def tri(v0, v1, v2, mat=0):
    t = np.zeros((), O.TRIANGLE)
    t["v0"][:3], t["v1"][:3], t["v2"][:3] = v0, v1, v2
    t["materialIndex"] = mat
    return t


def moller_trumbore(v0, v1, v2, o, d):
    """Textbook Moller-Trumbore in float64 (independent of the plane-based reference formula)."""
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-15:
        return None
    s = o - v0
    u = (s @ p) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    t = (e2 @ q) / det
    return t, u, v


def test_triangle_hit_matches_fp64_on_random_pairs():
    rng = np.random.default_rng(0)
    n_hit = 0
    for _ in range(4000):
        v = rng.uniform(-1, 1, (3, 3)).astype(np.float32)
        o = rng.uniform(-3, 3, 3).astype(np.float32)
        target = (v[0] + v[1] + v[2]) / 3 + rng.normal(0, 0.3, 3)
        d = (target - o).astype(np.float32)
        d /= np.linalg.norm(d)
        d = d.astype(np.float32)
        ok, t, u, w, nrm, bf = O.triangle_hit(tri(v[0], v[1], v[2]), o, d, 0.001, 1e7)
        ref = moller_trumbore(*(x.astype(np.float64) for x in (v[0], v[1], v[2], o, d)))
        if ref is None:
            continue
        rt, ru, rv = ref
        inside = ru > 1e-4 and rv > 1e-4 and ru + rv < 1 - 1e-4 and rt > 0.01
        outside = ru < -1e-4 or rv < -1e-4 or ru + rv > 1 + 1e-4 or rt < 0.0009
        n = np.cross(v[1].astype(np.float64) - v[0], v[2].astype(np.float64) - v[0])
        grazing = abs(n @ d.astype(np.float64)) / np.linalg.norm(n) < 2e-4
        if grazing:
            continue
        if inside:
            assert ok
            n_hit += 1
            assert abs(t - rt) <= 1e-5 * max(1.0, abs(rt))
            assert abs(u - ru) <= 1e-5 and abs(w - rv) <= 1e-5
            nn = n / np.linalg.norm(n)
            assert bf == (1 if nn @ d > 0 else 0)
            assert np.allclose(nrm, nn * (1 - 2 * bf), atol=1e-6)  # face-forwarded normal
        if outside:
            assert not ok
    assert n_hit > 500


def test_triangle_analytic_and_edge_cases():
    t0 = tri((0, 0, 2), (1, 0, 2), (0, 1, 2))
    ok, t, u, v, n, bf = O.triangle_hit(t0, (0.25, 0.25, 0), (0, 0, 1), 0.001, 1e7)
    assert ok and t == 2.0 and u == 0.25 and v == 0.25 and bf == 1 and tuple(n) == (0, 0, -1)
    assert O.triangle_hit(t0, (0, 0, 0), (0, 0, 1), 0.001, 1e7)[0]          # vertex v0: edges inclusive
    assert O.triangle_hit(t0, (0.5, 0.5, 0), (0, 0, 1), 0.001, 1e7)[0]      # on the hypotenuse a+b == 1
    assert not O.triangle_hit(t0, (0.6, 0.5, 0), (0, 0, 1), 0.001, 1e7)[0]
    assert O.triangle_hit(t0, (0.25, 0.25, 0), (0, 0, 1), 2.0, 2.0)[0]      # t == tmin == tmax accepted
    assert not O.triangle_hit(t0, (0.25, 0.25, 0), (0, 0, 1), 2.001, 1e7)[0]
    assert not O.triangle_hit(t0, (0.25, 0.25, 0), (1, 0, 0), 0.001, 1e7)[0]  # parallel: |n.d| < 1e-4
    d = np.array([1.0, 0.0, 0.00009], np.float32)
    assert not O.triangle_hit(t0, (0.25, 0.25, 1.9), d, 0.001, 1e7)[0]       # just under the cutoff
    assert not O.triangle_hit(tri((0, 0, 0), (1, 1, 1), (2, 2, 2)), (0.5, 0, -1), (0, 0, 1), 0.001, 1e7)[0]


def test_sphere_hit_analytic():
    s = np.zeros((), O.SPHERE)
    s["center"][:3] = (0, 0, 0)
    s["radius"] = 1.0
    ok, t, n, bf = O.sphere_hit(s, (0, 0, -5), (0, 0, 1), 0.001, 1e7)
    assert ok and t == 4.0 and tuple(n) == (0, 0, -1) and bf == 0
    ok, t, n, bf = O.sphere_hit(s, (0, 0, 0), (0, 0, 1), 0.001, 1e7)
    assert ok and t == 1.0 and bf == 1 and tuple(n) == (0, 0, -1)   # near root behind: far root, back face
    assert not O.sphere_hit(s, (0, 2, -5), (0, 0, 1), 0.001, 1e7)[0]
    assert not O.sphere_hit(s, (0, 0, -5), (0, 0, 1), 0.001, 3.9)[0]


def test_reference_slab_test():
    lo, hi = (-1, -1, -1), (1, 1, 1)
    assert O.aabb_hit_ref((0, 0, -5), (0, 0, 1), lo, hi)
    assert not O.aabb_hit_ref((0, 2, -5), (0, 0, 1), lo, hi)
    assert O.aabb_hit_ref((0, 0, -5), (0.1, 0.1, 1), lo, hi)
    assert O.aabb_hit_ref((0, 0, -5), (0, 0, -1), lo, hi)      # no t-interval: the line is tested (A9)
    assert not O.aabb_hit_ref((-3, -1, 0), (1, 1, 0), lo, hi)  # touches the corner only: tNear == tFar == 2
    assert O.aabb_hit_ref((-3, -1.1, 0), (1, 1, 0), lo, hi)    # crosses the corner region: [2, 2.1]


def test_camera_matches_fp64_pinhole():
    p = np.zeros((), O.PARAMS)
    p["camPos"][:3], p["camLookAt"][:3], p["camUpDir"][:3] = (1, 2, -6), (0, 0.5, 0), (0, 1, 0)
    p["verticalFOV"] = 40.0
    w, h = 321, 123
    cam = O.camera(p, w, h)
    pos, at, up = np.array([1, 2, -6.0]), np.array([0, 0.5, 0.0]), np.array([0, 1, 0.0])
    W = (pos - at) / np.linalg.norm(pos - at)
    U = np.cross(up, W); U /= np.linalg.norm(U)
    V = np.cross(W, U)
    vh = 2 * math.tan(math.radians(40) / 2) * 10.0
    vw = vh * w / h
    for (x, y) in [(0, 0), (w - 1, h - 1), (160, 61), (17, 100)]:
        target = pos - 10 * W + ((x + 0.5) / w - 0.5) * vw * U + (0.5 - (y + 0.5) / h) * vh * V
        dref = (target - pos) / np.linalg.norm(target - pos)
        o, d = O.primary_ray(cam, x, y)
        assert np.allclose(o, pos, atol=0) and np.allclose(d, dref, atol=2e-6)


def _rot(axis, a):
    c, s = math.cos(a), math.sin(a)
    if axis == "x":
        return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    if axis == "y":
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def test_transform_mat4_is_T_Ry_Rx_Rz_S():
    rng = np.random.default_rng(1)
    for _ in range(50):
        tr, sc, rot = rng.normal(size=3), rng.uniform(0.2, 3, 3), rng.uniform(-4, 4, 3)
        m = O.transform_mat4(tr, sc, rot).reshape(4, 4).T  # column-major storage -> row-major matrix
        ref = np.eye(4)
        ref[:3, :3] = _rot("y", rot[1]) @ _rot("x", rot[0]) @ _rot("z", rot[2]) @ np.diag(sc)
        ref[:3, 3] = tr
        assert np.allclose(m, ref, atol=2e-6 * max(1, np.abs(sc).max()))


def test_model_to_world():
    t = np.zeros(2, O.TRIANGLE)
    t["v0"][:, :3] = [(1, 0, 0), (0, 0, 0)]
    t["v1"][:, :3] = [(0, 1, 0), (1, 1, 1)]
    t["v2"][:, :3] = [(0, 0, 1), (2, 0, 0)]
    t["modelIndex"] = [0, 1]
    s = np.zeros(1, O.SPHERE)
    s["center"][0, :3] = (1, 1, 1)
    s["radius"] = 2.0
    s["modelIndex"] = 1
    models = np.stack([O.transform_mat4((0, 0, 0), (1, 1, 1), (0, 0, 0)),
                       O.transform_mat4((10, 20, 30), (1, 1, 1), (0, 0, 0))])
    wt, ws = O.model_to_world(t, s, models)
    assert np.array_equal(wt["v0"][0, :3], t["v0"][0, :3]) and wt["v0"][0, 3] == 1.0
    assert np.array_equal(wt["v1"][1, :3], [11, 21, 31])
    assert np.array_equal(ws["center"][0, :3], [11, 21, 31]) and ws["radius"][0] == 2.0  # radius unscaled (A17)
