"""Pins for the oracle's path tracer (raytraceBVH.comp:276-375), accumulation chain (RaytracerBVH.cpp:772,
1025-1050) and resolve (SingleTriangleFullScreen.frag:13-21), by closed-form scenes."""
import numpy as np

import oracle as O
from raytracergpu_mastersproject_amd import layouts as L
from raytracergpu_mastersproject_amd import workloads as W


def _emitter_wall_scene(albedo=(2.0, 3.0, 0.5)):
    """A huge emissive quad facing the camera covers every primary ray."""
    wall = W.make_triangles(np.array([(-100, -100, 5), (-100, -100, 5)]), np.array([(100, -100, 5), (100, 100, 5)]),
                            np.array([(100, 100, 5), (-100, 100, 5)]), 0)
    mats = W.make_materials([albedo], L.LIGHT)
    p = L.make_params((0, 0, 0), (0, 0, 1), (0, 1, 0), 40.0, 2, 0, 1, 8, 7)
    return wall, mats, p


def test_emitter_seen_directly_accumulates_albedo():
    wall, mats, p = _emitter_wall_scene()
    sc = O.OracleScene(wall, None, mats)
    acc, segs = sc.render(p, 8, 6, 4)
    assert segs == 8 * 6 * 4  # one segment per sample: LIGHT does not scatter
    assert np.array_equal(acc[:, :3], np.tile(np.float32([8.0, 12.0, 2.0]), (48, 1)))


def test_alpha_carries_the_rng_chain():
    """Per sample: state = seed0 + uint(alpha * 4294967294.0f); nextRandom = random(); alpha <- nextRandom.
    The first sample sees the clear value 1.0, whose product is 2^32 and saturates (A2)."""
    wall, mats, p = _emitter_wall_scene()
    sc = O.OracleScene(wall, None, mats)
    acc, _ = sc.render(p, 5, 3, 3)
    for (x, y) in [(0, 0), (4, 2), (2, 1)]:
        seed0 = O.rng_seed(x, y, int(p["randomState"]))
        alpha = np.float32(1.0)
        for _ in range(3):
            st = (seed0 + O.lib.orc_f2u_sat(float(alpha * np.float32(4294967294.0)))) & 0xFFFFFFFF
            vals, _ = O.rng_sequence(st, 1)
            alpha = vals[0]
        assert acc[y * 5 + x, 3] == alpha


def test_miss_everything_and_depth_zero_give_black():
    wall, mats, p = _emitter_wall_scene()
    p2 = p.copy()
    p2["camLookAt"][:3] = (0, 0, -1)  # look away from the wall
    sc = O.OracleScene(wall, None, mats)
    acc, segs = sc.render(p2, 4, 4, 2)
    assert np.all(acc[:, :3] == 0) and segs == 32
    p3 = p.copy()
    p3["maxRayTraceDepth"] = 0
    acc, segs = sc.render(p3, 4, 4, 2)
    assert np.all(acc[:, :3] == 0) and segs == 0


def test_diffuse_bounce_into_emitter():
    """Camera looks at a diffuse floor under an emissive ceiling spanning everything: every bounce off
    the floor that goes up hits the ceiling, so pixel = albedo_floor * emission when the scattered
    direction points upward (normal + unit vector is in the upper half-space)."""
    floor = W.quad(0.0, -1000, 1000, -1000, 1000, 0)
    ceiling = W.quad(10.0, -1000, 1000, -1000, 1000, 1, flip=True)
    tris = np.concatenate([floor, ceiling])
    mats = np.concatenate([W.make_materials([0.5, 0.25, 1.0], L.DIFFUSE), W.make_materials([2, 2, 2], L.LIGHT)])
    p = L.make_params((0, 1, 0), (0, 0, 1), (0, 1, 0), 40.0, 4, 0, 2, 8, 3)  # every primary ray goes down
    sc = O.OracleScene(tris, None, mats)
    acc, _ = sc.render(p, 6, 6, 1)
    vals = np.unique(np.round(acc[:, :3], 6), axis=0)
    for v in vals:
        assert np.allclose(v, [1.0, 0.5, 2.0]) or np.allclose(v, 0.0)
    assert np.mean(np.all(np.isclose(acc[:, :3], [1.0, 0.5, 2.0]), axis=1)) > 0.5


def test_resolve_closed_form():
    acc = np.array([[4.0, 1.0, 0.0, 0.3], [100.0, -1.0, 0.25, 1.0]], np.float32)
    out = O.resolve(acc, 4)
    assert np.allclose(out[0], [1.0, 0.5, 0.0, 1.0])
    assert out[1, 0] == 1.0 and np.isnan(out[1, 1]) and out[1, 2] == np.float32(0.25) and out[1, 3] == 1.0
