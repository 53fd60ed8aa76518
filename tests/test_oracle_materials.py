"""Closed-form pins for the oracle's METALLIC / DIELECTRIC scatter (the A5 extension: the reference itself
absorbs both, raytraceBVH.comp:100-115) and its composition inside the path tracer (:276-315).

Each scene's expected image is derived from geometry in float64 and the pinned RNG draws
(random.glsl:10-31, pinned in test_oracle_rng_trig.py), not from the oracle's scatter code:
- a fuzz-0 metal floor mirrors camera rays into an emitter that covers part of the sky: a pixel is
  albedo x emission exactly where the float64 mirror direction meets the emitter, else black;
- an ior-1.5 slab seen at normal incidence in front of an emitter: a pixel is the emission exactly
  when the chain of Schlick decisions (r0 = 0.04 at normal incidence) refracts in at the top, then
  out at the bottom, possibly after internal reflections, within the depth limit;
- rays that start inside a glass slab: beyond the critical angle (sin > 1/1.5) they are totally
  reflected back and forth and never reach the emitter above; below it the first refraction decides
  (out at the top: lit; out at the bottom: black);
- orc_scatter on its own: unit-length outputs, the mirror and Snell directions, the fuzz cone,
  the number of random draws each material consumes.
tests/test_oracle_mutants.py rebuilds the oracle with planted mistakes (refraction ratio chosen with
the face flag inverted, fuzz added after the normalisation, Schlick fed the wrong-signed cosine) and
checks that this file fails on each.
"""
import math

import numpy as np

import oracle as O
from raytracergpu_mastersproject_amd import layouts as L
from raytracergpu_mastersproject_amd import workloads as W

R0 = 0.04  # ((1 - 1.5) / (1 + 1.5))^2


def _draws(x, y, random_state, k):
    """The first sample's draws of pixel (x, y): [carry, then the scatters' draws in order]."""
    seed0 = O.rng_seed(x, y, random_state)
    st = (seed0 + O.lib.orc_f2u_sat(float(np.float32(1.0) * np.float32(4294967294.0)))) & 0xFFFFFFFF
    vals, _ = O.rng_sequence(st, k)
    return vals.astype(np.float64)


def _primary(params, w, h):
    cam = O.camera(params, w, h)
    out = {}
    for y in range(h):
        for x in range(w):
            o, d = O.primary_ray(cam, x, y)
            d64 = d.astype(np.float64)
            out[(x, y)] = (o.astype(np.float64), d64 / np.linalg.norm(d64))
    return out


def _render(tris, mats, params, w, h, spp=1):
    sc = O.OracleScene(tris, None, mats)
    acc, _ = sc.render(params, w, h, spp)
    return acc.reshape(h, w, 4)


def _schlick(cos, ri):
    r0 = ((1 - ri) / (1 + ri)) ** 2
    return r0 + (1 - r0) * (1 - cos) ** 5


def test_fuzz0_metal_mirror_into_an_emitter():
    floor = W.quad(0.0, -50, 50, -50, 50, 0, flip=True)
    light = W.quad(5.0, -3, 3, 0, 6, 1)
    tris = np.concatenate([floor, light])
    albedo, emission = np.float32([0.9, 0.6, 0.3]), np.float32([4, 2, 1])
    mats = np.concatenate([W.make_materials([albedo], L.METALLIC, 0.0), W.make_materials([emission], L.LIGHT)])
    w, h = 48, 36
    p = L.make_params((0, 3, -6), (0, 0, 0), (0, 1, 0), 60.0, len(tris), 0, 2, 8, 11)
    img = _render(tris, mats, p, w, h)
    lit = dark = 0
    for (x, y), (o, d) in _primary(p, w, h).items():
        got = img[y, x, :3]
        if d[1] >= 0:
            continue  # above the horizon (the emitter is not in view here)
        t = -o[1] / d[1]
        hit = o + t * d
        if abs(hit[0]) > 49 or abs(hit[2]) > 49:
            continue
        m = d * np.array([1, -1, 1])  # mirror about the floor's normal (0, 1, 0)
        s = (5.0 - hit[1]) / m[1]
        q = hit + s * m
        margin = min(abs(q[0] + 3), abs(q[0] - 3), abs(q[2]), abs(q[2] - 6))
        if margin < 1e-3:
            continue
        if -3 < q[0] < 3 and 0 < q[2] < 6:
            assert np.array_equal(got, emission * albedo), (x, y, got)
            lit += 1
        else:
            assert np.array_equal(got, np.zeros(3, np.float32)), (x, y, got)
            dark += 1
    assert lit > 100 and dark > 100


def _slab_scene(ior=1.5):
    top = W.quad(1.2, -10, 10, -10, 10, 0, flip=True)  # outward normal +y
    bottom = W.quad(1.0, -10, 10, -10, 10, 0)  # outward normal -y
    light = W.quad(0.0, -10, 10, -10, 10, 1)
    tris = np.concatenate([top, bottom, light])
    emission = np.float32([3, 5, 7])
    mats = np.concatenate([W.make_materials([[1, 1, 1]], L.DIELECTRIC, ior), W.make_materials([emission], L.LIGHT)])
    return tris, mats, emission


def test_normal_incidence_through_a_glass_slab():
    tris, mats, emission = _slab_scene()
    w, h, depth, rs = 24, 24, 8, 5
    # looking straight down with a 0.01-degree field of view: every ray is within 1e-4 rad of the normal
    p = L.make_params((0, 20, 0), (0, 0, 0), (0, 0, 1), 0.01, len(tris), 0, 2, depth, rs)
    img = _render(tris, mats, p, w, h)
    lit = dark = 0
    for y in range(h):
        for x in range(w):
            u = _draws(x, y, rs, depth + 2)
            used, seg, on = [], 1, True  # segment 1: camera -> top face (entering)
            if u[1] < R0:
                on = False  # reflected back to the sky
            used.append(u[1])
            j = 2
            while on:
                seg += 1  # inside, at the bottom face
                used.append(u[j])
                if u[j] >= R0:
                    on = seg + 1 <= depth  # out through the bottom into the emitter
                    break
                seg += 1  # reflected up, at the top face from inside
                if seg > depth:
                    on = False
                    break
                used.append(u[j + 1])
                if u[j + 1] >= R0:
                    on = False  # out through the top to the sky
                    break
                j += 2
                if seg + 1 > depth:
                    on = False
                    break
            if min(abs(v - R0) for v in used) < 1e-5:
                continue
            want = emission if on else np.zeros(3, np.float32)
            assert np.array_equal(img[y, x, :3], want), (x, y, img[y, x, :3], used)
            lit += on
            dark += not on
    assert lit > 400 and dark > 10


def test_total_internal_reflection_inside_glass():
    top = W.quad(1.0, -1000, 1000, -1000, 1000, 0, flip=True)  # outward +y
    bottom = W.quad(0.0, -1000, 1000, -1000, 1000, 0)  # outward -y
    light = W.quad(3.0, -1e4, 1e4, -1e4, 1e4, 1)
    tris = np.concatenate([top, bottom, light])
    emission = np.float32([2, 2, 2])
    mats = np.concatenate([W.make_materials([[1, 1, 1]], L.DIELECTRIC, 1.5), W.make_materials([emission], L.LIGHT)])
    w, h, depth, rs = 40, 40, 8, 9
    c = math.sqrt(0.5)
    p = L.make_params((0, 0.5, 0), (0, 0.5 + c, c), (0, 1, 0), 90.0, len(tris), 0, 2, depth, rs)
    img = _render(tris, mats, p, w, h)
    n_tir = n_ref = 0
    for (x, y), (o, d) in _primary(p, w, h).items():
        if d[1] < 0.3:
            continue
        cos = d[1]
        sin = math.sqrt(1 - cos * cos)
        got = img[y, x, :3]
        if sin > 0.70:  # beyond the critical angle: reflected at every face, never out
            assert np.array_equal(got, np.zeros(3, np.float32)), (x, y, sin, got)
            n_tir += 1
        elif sin < 0.62:
            r = _schlick(cos, 1.5)
            u = _draws(x, y, rs, depth + 2)
            on, used = False, []
            for k in range(1, depth + 1):  # visit k: top face when k is odd, bottom when even
                used.append(u[k])
                if u[k] >= r:
                    on = (k % 2 == 1) and k + 1 <= depth
                    break
            if min(abs(v - r) for v in used) < 1e-5:
                continue
            assert np.array_equal(got, emission if on else np.zeros(3, np.float32)), (x, y, got, used, r)
            n_ref += 1
    assert n_tir > 100 and n_ref > 100


def _unit(rng, n=3):
    v = rng.normal(size=n)
    return v / np.linalg.norm(v)


def _steps(state, k):
    for _ in range(k):
        state = O.lib.orc_rng_step(state)
    return state


def test_metal_scatter_directions():
    rng = np.random.default_rng(5)
    for fuzz in (0.0, 0.2, 0.5):
        m = W.make_materials([[0.8, 0.5, 0.25]], L.METALLIC, fuzz)[0]
        for _ in range(300):
            n = _unit(rng)
            d = _unit(rng)
            if d @ n > -0.05:
                d = d - 2 * (d @ n) * n if d @ n > 0 else d - 0.2 * n
                d /= np.linalg.norm(d)
            n32, d32 = n.astype(np.float32), d.astype(np.float32)
            state = int(rng.integers(0, 2**32))
            ok, att, out, st = O.scatter(m, d32, n32, False, state)
            assert st == _steps(state, 3)  # three draws: one point in the unit sphere
            assert np.array_equal(att, np.float32([0.8, 0.5, 0.25]))
            o64 = out.astype(np.float64)
            assert abs(np.linalg.norm(o64) - 1.0) < 1e-6
            dn, nn = d32.astype(np.float64), n32.astype(np.float64)
            mirror = dn - 2 * (dn @ nn) * nn
            mirror /= np.linalg.norm(mirror)
            ang = math.atan2(np.linalg.norm(np.cross(o64, mirror)), o64 @ mirror)
            assert ang <= math.asin(min(fuzz, 1.0)) + 1e-5, (fuzz, ang)
            assert ok == bool(o64 @ nn > 0)
        if fuzz == 0.0:
            assert ang < 1e-5


def test_dielectric_scatter_directions_and_draws():
    rng = np.random.default_rng(6)
    m = W.make_materials([[1, 1, 1]], L.DIELECTRIC, 1.5)[0]
    n_refr = n_refl = 0
    for _ in range(600):
        n = _unit(rng)
        d = _unit(rng)
        if d @ n > 0:
            d = -d
        backface = bool(rng.integers(0, 2))
        n32, d32 = n.astype(np.float32), d.astype(np.float32)
        dn, nn = d32.astype(np.float64), n32.astype(np.float64)
        dn /= np.linalg.norm(dn)
        nn /= np.linalg.norm(nn)
        cos = -(dn @ nn)
        sin = math.sqrt(max(0.0, 1 - cos * cos))
        ri = 1.5 if backface else 1 / 1.5  # leaving the medium through a back face, entering through a front face
        state = int(rng.integers(0, 2**32))
        ok, att, out, st = O.scatter(m, d32, n32, backface, state)
        assert ok and np.array_equal(att, np.ones(3, np.float32))
        o64 = out.astype(np.float64)
        assert abs(np.linalg.norm(o64) - 1.0) < 1e-6
        mirror = dn + 2 * cos * nn
        if abs(ri * sin - 1) < 1e-4:
            continue
        if ri * sin > 1:  # total internal reflection: no random draw
            assert st == state
            assert np.linalg.norm(o64 - mirror) < 1e-5
            n_refl += 1
            continue
        assert st == _steps(state, 1)
        u, _ = O.rng_sequence(state, 1)
        r = _schlick(cos, ri)
        if abs(float(u[0]) - r) < 1e-5:
            continue
        if u[0] < r:
            assert np.linalg.norm(o64 - mirror) < 1e-5
            n_refl += 1
        else:  # Snell: sin(t) = ri sin(i), continuing through the surface
            assert o64 @ nn < 0
            sin_t = np.linalg.norm(np.cross(o64, nn))
            assert abs(sin_t - ri * sin) < 1e-5, (sin_t, ri * sin)
            n_refr += 1
    assert n_refr > 100 and n_refl > 20


def test_diffuse_light_and_reference_materials():
    rng = np.random.default_rng(7)
    n = _unit(rng).astype(np.float32)
    d = (-n).astype(np.float32)
    diffuse = W.make_materials([[0.5, 0.5, 0.5]], L.DIFFUSE)[0]
    ok, att, out, st = O.scatter(diffuse, d, n, False, 123)
    assert ok and st == _steps(123, 3) and abs(np.linalg.norm(out) - 1) < 1e-6 and out @ n >= 0
    light = W.make_materials([[1, 1, 1]], L.LIGHT)[0]
    ok, _, _, st = O.scatter(light, d, n, False, 123)
    assert not ok and st == 123
    # the reference's own behaviour (A5): METALLIC and DIELECTRIC absorb, drawing nothing
    for mt, extra in ((L.METALLIC, 0.3), (L.DIELECTRIC, 1.5)):
        mm = W.make_materials([[1, 1, 1]], mt, extra)[0]
        ok, _, _, st = O.scatter(mm, d, n, False, 123, flags=O.RENDER_REFERENCE_MATERIALS)
        assert not ok and st == 123
