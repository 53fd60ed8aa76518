"""The two benchmark workloads at their full sizes, checked where the oracle cannot run every case.

Ray query: 4,194,304 clustered triangles, 16,777,216 incoherent rays (the bench's headline).
- The device builder's binary LBVH equals the oracle's builder byte for byte.
- On every ray: any-hit finds a hit exactly when closest-hit does; a closest hit lies in
  [tmin, tmax], names a real triangle and is no farther than the any-hit; two launch shapes
  give bitwise-identical records.
- On a strided sample of 262,144 rays: hits equal the oracle's reference traversal (the
  reference's right-first hitBVH over its binary LBVH) bit for bit, and the kernel's per-ray
  AABB and triangle test counts equal the oracle's replay of the same wide traversal.

Fused many-scene ray query: 256 scenes (65.6 M triangles), 55.1 M rays in batches of 4,096 to
1,048,576.
- The fused launch's hits equal each scene's plain launch bit for bit, on every ray.
- On every 97th ray of four scenes (the smallest, the largest and two between), hits equal the
  oracle's reference traversal bit for bit.

Path tracing: the 1,003,522-triangle icosphere, 1920x1080, 16 spp, depth 8 (rendered bounce by
bounce by default at this size).
- The device builder's binary LBVH of the icosphere equals the oracle's builder byte for byte.
- Bands of rows (the middle of the sphere, and rows from 600 where the reference's seed
  formula aliases) equal the oracle's path tracer bit for bit, RGBA including the RNG carry.
- The whole image equals the single persistent launch's image: the schedule changes when a
  path's segments run, never what they compute.
"""
import numpy as np
import pytest

from raytracergpu_mastersproject_amd import layouts as L
from raytracergpu_mastersproject_amd import workloads as W

from helpers import compare_hits, upload_both

pytestmark = pytest.mark.gpu


def test_ray_query_at_full_size(ctx, native, O):
    tris, mats, rays = W.ray_query_workload()
    n = len(rays)
    print(f"full-size ray query: {len(tris)} triangles, {n} rays", flush=True)
    sid, osc = upload_both(ctx, O, tris, None, mats)
    nodes, _, slots, _, _ = O.build_lbvh(tris)
    assert np.array_equal(ctx.nodes(sid).view(np.uint8), np.ascontiguousarray(nodes).view(np.uint8))
    assert np.array_equal(ctx.slots(sid), slots)

    d_rays = native.DeviceArray.from_numpy(rays)
    d_hits = native.DeviceArray(16 * n)
    d_counts = native.DeviceArray(8 * n)
    ctx.trace(sid, d_rays, n, d_hits, d_counts)
    g = d_hits.to_numpy(L.HIT, n)
    counts = d_counts.to_numpy(np.uint32, 2 * n).reshape(n, 2)
    ctx.trace(sid, d_rays, n, d_hits, cfg=native.launch_config(64, 8, 1, 32))
    g2 = d_hits.to_numpy(L.HIT, n)
    ctx.trace(sid, d_rays, n, d_hits, any_hit=True)
    a = d_hits.to_numpy(L.HIT, n)
    del d_rays, d_hits, d_counts

    assert np.array_equal(g.view(np.uint32), g2.view(np.uint32))
    hit = g["prim"] != L.NO_HIT
    assert np.array_equal(hit, a["prim"] != L.NO_HIT)
    assert 0.1 < hit.mean() < 0.9
    assert np.all(g["prim"][hit] < len(tris)) and np.all(a["prim"][hit] < len(tris))
    assert np.all(g["t"][hit] >= rays["tmin"][hit]) and np.all(g["t"][hit] <= rays["tmax"][hit])
    assert np.all(g["t"][hit] <= a["t"][hit])

    sample = np.arange(0, n, 64)
    r = np.ascontiguousarray(rays[sample])
    assert compare_hits(g[sample], osc.trace(r, O.REF_BVH)) == 0
    ow, oc = osc.trace(r, O.WIDE, counts=True)
    assert compare_hits(g[sample], ow) == 0
    assert np.array_equal(counts[sample, 0], oc["n_box"]) and np.array_equal(counts[sample, 1], oc["n_tri"])
    ctx.release(sid)


def test_fused_many_scene_query_at_full_size(ctx, native, O):
    work = W.batched_workload()
    print(f"full-size fused query: {len(work)} scenes, {sum(len(t) for t, _, _ in work)} triangles, "
          f"{sum(len(r) for _, _, r in work)} rays", flush=True)
    items = []
    for tris, mats, rays in work:
        sid = ctx.upload(tris, None, mats)
        items.append((sid, len(rays), native.DeviceArray.from_numpy(rays), native.DeviceArray(16 * len(rays))))
    ctx.trace_batch(items)
    fused = [dh.to_numpy(L.HIT, n) for (_, n, _, dh) in items]
    for (sid, n, dr, dh), f in zip(items, fused):
        ctx.trace(sid, dr, n, dh)
        assert np.array_equal(dh.to_numpy(L.HIT, n).view(np.uint32), f.view(np.uint32)), sid
    sizes = np.array([len(t) for t, _, _ in work])
    order = np.argsort(sizes, kind="stable")
    for k in (order[0], order[len(order) // 3], order[2 * len(order) // 3], order[-1]):
        tris, mats, rays = work[k]
        nodes, _, slots, _, _ = O.build_lbvh(tris)
        osc = O.OracleScene(tris, None, mats, nodes, slots)
        sample = np.arange(0, len(rays), 97)
        assert compare_hits(fused[k][sample], osc.trace(np.ascontiguousarray(rays[sample]), O.REF_BVH)) == 0, k
    assert 0.01 < np.mean(np.concatenate([f["prim"] != L.NO_HIT for f in fused])) < 0.5  # 3.3 % of these rays hit
    for (sid, _, _, _) in items:
        ctx.release(sid)


def _render(native, ctx, sid, s, flags=0):
    w, h = s["width"], s["height"]
    acc = native.DeviceArray(16 * w * h)
    ctx.accum_clear(acc, w * h)
    ctx.render(sid, s["params"], w, h, s["spp"], acc, flags=flags)
    return acc.to_numpy(np.float32, 4 * w * h).reshape(h, w, 4)


def test_path_traced_frame_at_full_size(ctx, native, O):
    s = W.icosphere_scene()
    w, h, spp = s["width"], s["height"], s["spp"]
    assert (w, h, spp, int(s["params"]["maxRayTraceDepth"])) == (1920, 1080, 16, 8)
    print(f"full-size frame: {len(s['tris'])} triangles, {w}x{h}, {spp} spp", flush=True)
    sid, osc = upload_both(ctx, O, s["tris"], None, s["materials"])
    nodes, _, slots, _, _ = O.build_lbvh(s["tris"])
    assert np.array_equal(ctx.nodes(sid).view(np.uint8), np.ascontiguousarray(nodes).view(np.uint8))
    assert np.array_equal(ctx.slots(sid), slots)
    g = _render(native, ctx, sid, s)
    assert np.all(np.isfinite(g)) and np.all(g[..., :3] >= 0.0)
    lit = g[..., :3].sum(axis=2) > 0.0
    assert 0.05 < lit.mean() < 0.95

    acc = None
    for r0, r1 in ((520, 552), (600, 632)):
        acc, _ = osc.render(s["params"], w, h, spp, accum=acc, method=O.WIDE, rows=(r0, r1))
    o = acc.reshape(h, w, 4)
    for r0, r1 in ((520, 552), (600, 632)):
        assert np.array_equal(g[r0:r1].view(np.uint32), o[r0:r1].view(np.uint32)), (r0, r1)

    gp = _render(native, ctx, sid, s, native.RENDER_PERSISTENT_PATHS)
    assert np.array_equal(g.view(np.uint32), gp.view(np.uint32))
    ctx.release(sid)
