"""Pins for the oracle's LBVH builder (GetEnclosingAABB.comp, GenerateMortonCodesOfPrimitives.comp,
RadixSortSimple.comp, ConstructHLBVH.comp, ConstructAABBsOfInternalNodes.comp) and for its traversals
(hitBVH raytraceBVH.comp:195-265 vs the brute force of raytrace.comp:167-190)."""
import numpy as np
import pytest

import oracle as O
from raytracergpu_mastersproject_amd import workloads as W


def tris_at(centroids, r=0.0001):
    n = len(centroids)
    t = np.zeros(n, O.TRIANGLE)
    c = np.asarray(centroids, np.float64)
    t["v0"][:, :3] = c + [r, 0, 0]
    t["v1"][:, :3] = c + [0, r, 0]
    t["v2"][:, :3] = c + [-r, -r, 0]
    return t


def interleave(qx, qy, qz):
    """Morton code by explicit bit interleaving: bit 3i = x_i, 3i+1 = y_i, 3i+2 = z_i."""
    code = 0
    for i in range(10):
        code |= ((qx >> i) & 1) << (3 * i) | ((qy >> i) & 1) << (3 * i + 1) | ((qz >> i) & 1) << (3 * i + 2)
    return code


def test_morton_matches_bit_interleave():
    rng = np.random.default_rng(0)
    q = rng.integers(0, 1024, (500, 3))
    c = (q + 0.5) / 1024.0  # quantises back to q with emin 0, span 1
    t = np.zeros(len(c), O.TRIANGLE)
    for k in ("v0", "v1", "v2"):
        t[k][:, :3] = c
    codes = O.morton_codes(t, np.zeros(0, O.SPHERE), (0, 0, 0), (1, 1, 1))
    for (qx, qy, qz), code in zip(q, codes):
        assert code == interleave(qx, qy, qz)
    # coordinate exactly at the top of the box quantises to 1024 and is clamped to 1023
    t2 = t[:1].copy()
    for k in ("v0", "v1", "v2"):
        t2[k][:, :3] = (1.0, 1.0, 1.0)
    assert O.morton_codes(t2, np.zeros(0, O.SPHERE), (0, 0, 0), (1, 1, 1))[0] == interleave(1023, 1023, 1023)


def test_reference_morton_quantisation_quirk():
    """A12: the reference divides coord by span without subtracting the minimum; negative
    coordinates saturate to 0 under the AMD conversion rule."""
    t = np.zeros(2, O.TRIANGLE)
    for k in ("v0", "v1", "v2"):
        t[k][:, :3] = [(0.25, 0.5, 0.75), (-0.5, 0.25, 0.0)]
    c = O.morton_codes(t, np.zeros(0, O.SPHERE), (-1, -1, -1), (1, 1, 1), flags=O.BUILD_REFERENCE_MORTON)
    assert c[0] == interleave(128, 256, 384)   # 0.25/2*1024 = 128 ...
    assert c[1] == interleave(0, 128, 0)


def test_enclosing_aabb_padding():
    t = tris_at([(0, 0, 0), (1, 0, 0)], r=0.0)
    mn, mx = O.enclosing_aabb(t, np.zeros(0, O.SPHERE))
    assert mn[0] == 0 and mx[0] == 1
    assert mn[1] == np.float32(-0.0005) and mx[1] == np.float32(0.0005)  # flat extent padded by DELTA/2


def test_stable_radix_sort_matches_python_stable_sort():
    rng = np.random.default_rng(1)
    keys = rng.integers(0, 50, 10000).astype(np.uint32) * np.uint32(2 ** 20) + rng.integers(0, 3, 10000).astype(
        np.uint32)
    vals = np.arange(10000, dtype=np.uint32)
    k, v = O.stable_radix_sort(keys, vals)
    order = sorted(range(10000), key=lambda i: int(keys[i]))
    assert np.array_equal(v, np.array(order, np.uint32))
    assert np.array_equal(k, keys[order])


def test_hand_built_four_primitive_tree():
    """Sorted codes along x at 0.1, 0.3, 0.6, 0.9: Karras gives root [0,3] split 1, node 1 = leaves 0,1,
    node 2 = leaves 2,3 (leaf k at index N-1+k = 3+k)."""
    t = tris_at([(0.9, 0.5, 0.5), (0.1, 0.5, 0.5), (0.6, 0.5, 0.5), (0.3, 0.5, 0.5)])
    nodes, parent, slots, codes, prims = O.build_lbvh(t)
    assert list(prims) == [1, 3, 2, 0] and list(slots) == [1, 3, 2, 0]
    assert (nodes[0]["left"], nodes[0]["right"]) == (1, 2)
    assert (nodes[1]["left"], nodes[1]["right"]) == (3, 4)
    assert (nodes[2]["left"], nodes[2]["right"]) == (5, 6)
    assert list(parent[1:]) == [0, 0, 1, 1, 2, 2]
    # reference leaf placement (A11): leaf 3+k holds primitive k
    nodes_r, _, slots_r, _, _ = O.build_lbvh(t, flags=O.BUILD_REFERENCE_LEAVES)
    assert list(slots_r) == [0, 1, 2, 3] and list(nodes_r[3:]["primitiveIndex"]) == [0, 1, 2, 3]


def _check_tree(nodes, parent, slots, tris, spheres):
    n = len(slots)
    assert len(nodes) == 2 * n - 1
    seen = np.zeros(2 * n - 1, np.int32)
    rng_of = {}

    def box(i):
        nd = nodes[i]
        return np.array([nd["minX"], nd["minY"], nd["minZ"]]), np.array([nd["maxX"], nd["maxY"], nd["maxZ"]])

    stack = [0]
    seen[0] += 1
    order = []
    while stack:
        i = stack.pop()
        order.append(i)
        nd = nodes[i]
        if nd["left"] == 0 and nd["right"] == 0:
            assert i >= n - 1 or n == 1
            continue
        for c in (nd["left"], nd["right"]):
            seen[c] += 1
            assert parent[c] == i
            stack.append(c)
    assert np.all(seen == 1), "every node reachable exactly once"
    for i in reversed(order):
        nd = nodes[i]
        if nd["left"] == 0 and nd["right"] == 0:
            k = i - (n - 1)
            rng_of[i] = (k, k)
            p = slots[k]
            lo, hi = box(i)
            if p & O.SPHERE_BIT:
                s = spheres[int(p) & 0x7FFFFFFF]
                assert np.all(lo <= s["center"][:3] - s["radius"] + 1e-6)
            else:
                v = np.stack([tris[p]["v0"][:3], tris[p]["v1"][:3], tris[p]["v2"][:3]])
                assert np.all(lo <= v.min(0)) and np.all(hi >= v.max(0))
                assert np.all(hi - lo >= np.float32(0.001) - 1e-7)  # padded extent
            continue
        a, b = rng_of[nd["left"]], rng_of[nd["right"]]
        assert a[1] + 1 == b[0], "subtrees cover contiguous leaf ranges, left before right"
        rng_of[i] = (a[0], b[1])
        la, ha = box(nd["left"])
        lb, hb = box(nd["right"])
        lo, hi = box(i)
        assert np.array_equal(lo, np.minimum(la, lb)) and np.array_equal(hi, np.maximum(ha, hb))
    assert rng_of[0] == (0, n - 1)


@pytest.mark.parametrize("n", [1, 2, 5, 64, 3000])
def test_lbvh_invariants(n):
    tris, _ = W.cluster_scene(4, max(1, n // 4), seed=n) if n >= 4 else W.cluster_scene(1, n, seed=n)
    nodes, parent, slots, codes, prims = O.build_lbvh(tris)
    assert np.all(np.diff(codes.astype(np.int64)) >= 0)
    if len(tris) > 1:
        _check_tree(nodes, parent, slots, tris, np.zeros(0, O.SPHERE))


def test_lbvh_invariants_with_duplicates_and_spheres():
    tris, _ = W.cluster_scene(2, 500, seed=5)
    tris = np.concatenate([tris, tris[:200]])
    sph = np.zeros(4, O.SPHERE)
    sph["center"][:, :3] = [(0, 0, 0), (0.5, 0.5, 0.5), (0.5, 0.5, 0.5), (-1, 0, 1)]
    sph["radius"] = 0.1
    nodes, parent, slots, codes, prims = O.build_lbvh(tris, sph)
    _check_tree(nodes, parent, slots, tris, sph)
    assert np.sum((slots & O.SPHERE_BIT) != 0) == 4


def _brute_fp64(tris, rays):
    """Closest hit in float64 (Moller-Trumbore), ignoring exact ties."""
    v0 = np.stack([tris["v0"][:, :3], tris["v1"][:, :3], tris["v2"][:, :3]], 1).astype(np.float64)
    out = []
    for r in rays:
        o, d = r["o"].astype(np.float64), r["d"].astype(np.float64)
        e1, e2 = v0[:, 1] - v0[:, 0], v0[:, 2] - v0[:, 0]
        p = np.cross(d, e2)
        det = np.einsum("ij,ij->i", e1, p)
        with np.errstate(divide="ignore", invalid="ignore"):
            s = o - v0[:, 0]
            u = np.einsum("ij,ij->i", s, p) / det
            q = np.cross(s, e1)
            v = (q @ d) / det
            t = np.einsum("ij,ij->i", e2, q) / det
        ok = (u >= 0) & (v >= 0) & (u + v <= 1) & (t >= r["tmin"]) & (t <= r["tmax"]) & np.isfinite(t)
        out.append((np.argmin(np.where(ok, t, np.inf)), np.min(np.where(ok, t, np.inf))))
    return out


def test_closest_hit_brute_equals_reference_traversal_and_fp64():
    s = W.soup_scene()
    sc = O.OracleScene(s["tris"], None, s["materials"])
    lo, hi = W.scene_bounds(s["tris"])
    rays = W.incoherent_rays(3000, lo, hi, seed=17)
    hb = sc.trace(rays, O.BRUTE)
    hr = sc.trace(rays, O.REF_BVH)
    assert np.array_equal(hb["prim"], hr["prim"]) and np.array_equal(hb["t"].view(np.uint32), hr["t"].view(np.uint32))
    ref = _brute_fp64(s["tris"], rays[:600])
    for h, (k, t) in zip(hb[:600], ref):
        if np.isfinite(t):
            assert h["prim"] == k and abs(h["t"] - t) <= 1e-5 * max(1, t)
        else:
            assert h["prim"] == O.NO_HIT
    # any-hit existence equals closest-hit existence
    ha = sc.trace(rays, O.BRUTE, any_hit=True)
    assert np.array_equal(ha["prim"] != O.NO_HIT, hb["prim"] != O.NO_HIT)


def test_ties_choose_the_primitive_tested_last():
    """Duplicated triangles give equal t; the reference's right-first traversal accepts equal t, so the
    copy in the leftmost leaf slot wins; the reference-order brute force agrees."""
    tris, _ = W.cluster_scene(1, 300, seed=3)
    tris = np.concatenate([tris, tris])
    sc = O.OracleScene(tris)
    lo, hi = W.scene_bounds(tris)
    rays = W.incoherent_rays(3000, lo, hi, seed=4)
    hb, hr = sc.trace(rays, O.BRUTE), sc.trace(rays, O.REF_BVH)
    assert np.array_equal(hb["prim"], hr["prim"])
    slot_of = np.empty(len(tris), np.int64)
    slot_of[sc.slots] = np.arange(len(tris))
    hit = hb["prim"] != O.NO_HIT
    twin = np.where(hb["prim"][hit] < 300, hb["prim"][hit] + 300, hb["prim"][hit] - 300)
    assert np.all(slot_of[hb["prim"][hit]] < slot_of[twin])


def test_rays_in_a_box_face_plane():
    """A ray with a zero direction component lying exactly in the plane of a box face.

    The reference's AABBhitCheck (raytraceBVH.comp:184-193, restated with GLSL min/max) gets
    (face - origin) * inf = NaN on that axis and culls the box on either face; brute force reports
    the hit.  The wide traversal (replayed here over a hand-built two-leaf wide node) takes the zero
    as 2^-100 with the zero's sign: the ray moves into the box from one face, out of it from the
    other, so it enters the box on the low face with +0 and on the high face with -0, and is culled
    on the other two, where the nudged ray only touches the box at t = 0."""
    t = np.zeros(2, O.TRIANGLE)
    t["v0"][0, :3], t["v1"][0, :3], t["v2"][0, :3] = [0, -1, 1], [1, -1, 1], [0, 1, 1]      # x in [0, 1], z = 1
    t["v0"][1, :3], t["v1"][1, :3], t["v2"][1, :3] = [2, -1, 4], [3, -1, 4], [3, 1, 4]      # x in [2, 3], z = 4
    t["v0"][:, 3] = t["v1"][:, 3] = t["v2"][:, 3] = 1.0
    mats = W.make_materials([0.5, 0.5, 0.5], 1)
    width = 2
    node = np.zeros(8 * width, np.uint32)
    f = node.view(np.float32)
    for c in range(2):
        v = np.stack([t["v0"][c, :3], t["v1"][c, :3], t["v2"][c, :3]])
        lo, hi = v.min(0), v.max(0)
        for k in range(3):
            f[(2 * k) * width + c], f[(2 * k + 1) * width + c] = lo[k], hi[k]
        node[6 * width + c] = (1 << 31) | c  # leaf: one primitive at slot c
    wide_scene = O.OracleScene(t, None, mats, None, np.array([0, 1], np.uint32), node, width)
    ref_scene = O.OracleScene(t, None, mats)  # the reference LBVH over the same triangles
    r = np.zeros(5, O.RAY)
    r["tmin"], r["tmax"] = 0.001, 1e7
    # low face x = 0 of box 0 with both zero signs, high face x = 3 of box 1 with both; strictly inside
    r["o"] = [[0, 0, 0], [0, 0, 0], [3, 0.5, 0], [3, 0.5, 0], [0.5, 0, 0]]
    r["d"] = [[0, 0, 1], [-0.0, 0, 1], [0, 0, 1], [-0.0, 0, 1], [0, 0, 1]]
    brute = ref_scene.trace(r, O.BRUTE)
    assert list(brute["prim"]) == [0, 0, 1, 1, 0]
    ref = ref_scene.trace(r, O.REF_BVH)
    assert list(ref["prim"][:4]) == [O.NO_HIT] * 4  # culled in the face plane
    wide = wide_scene.trace(r, O.WIDE)
    assert list(wide["prim"][:4]) == [0, O.NO_HIT, O.NO_HIT, 1]  # entered where the nudged ray moves inward
    assert np.array_equal(wide["t"][[0, 3]], brute["t"][[0, 3]])
    for hits in (ref, wide):
        assert hits["prim"][4] == 0 and hits["t"][4] == brute["t"][4]  # strictly inside: found
