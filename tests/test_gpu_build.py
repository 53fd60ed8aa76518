"""Device LBVH builder vs the oracle's restatement of the reference build shaders (bit-exact)."""
import numpy as np
import pytest

from raytracergpu_mastersproject_amd import layouts as L
from raytracergpu_mastersproject_amd import workloads as W

from helpers import small_cluster_scene, with_spheres

pytestmark = pytest.mark.gpu


def _nodes_equal(a, b):
    return np.array_equal(a.view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_builder_output_matches_oracle_soup(ctx, native, O, flags):
    s = W.soup_scene()
    sid = ctx.upload(s["tris"], None, s["materials"], flags=flags | native.BUILD_KEEP_BINARY)
    gnodes, gslots = ctx.nodes(sid), ctx.slots(sid)
    onodes, _, oslots, _, _ = O.build_lbvh(s["tris"], flags=flags)
    assert len(gnodes) == 2 * len(s["tris"]) - 1
    assert _nodes_equal(gnodes, onodes), "binary node buffer differs from the reference restatement"
    assert np.array_equal(gslots, oslots)
    ctx.release(sid)


@pytest.mark.parametrize("n", [1, 2, 3, 7, 64, 1000, 4097])
def test_builder_small_sizes(ctx, native, O, n):
    tris, mats = W.cluster_scene(1, n, seed=n)
    sid = ctx.upload(tris, None, mats, flags=native.BUILD_KEEP_BINARY)
    onodes, _, oslots, _, _ = O.build_lbvh(tris)
    assert _nodes_equal(ctx.nodes(sid), onodes)
    assert np.array_equal(ctx.slots(sid), oslots)
    inf = ctx.info(sid)
    assert inf.nprim == n and inf.nwide >= 1
    ctx.release(sid)


def test_builder_clusters_with_duplicates(ctx, native, O):
    # many identical centroids exercise the Karras index tie-break (ConstructHLBVH.comp:64-68)
    tris, mats = small_cluster_scene(16, 4096)
    tris = np.concatenate([tris, tris[:3000]])
    sid = ctx.upload(tris, None, mats, flags=native.BUILD_KEEP_BINARY)
    onodes, _, oslots, _, _ = O.build_lbvh(tris)
    assert _nodes_equal(ctx.nodes(sid), onodes)
    assert np.array_equal(ctx.slots(sid), oslots)
    ctx.release(sid)


def test_builder_mixed_spheres(ctx, native, O):
    tris, mats = small_cluster_scene(4, 500)
    sph = with_spheres(tris, mats, 9)
    sid = ctx.upload(tris, sph, mats, flags=native.BUILD_KEEP_BINARY)
    onodes, _, oslots, _, _ = O.build_lbvh(tris, sph)
    assert _nodes_equal(ctx.nodes(sid), onodes)
    assert np.array_equal(ctx.slots(sid), oslots)
    assert np.sum((oslots & O.SPHERE_BIT) != 0) == 9
    ctx.release(sid)


def test_model_to_world_on_device(ctx, native, O):
    s = W.soup_scene(ntri=200)
    tris = s["tris"].copy()
    models = np.stack([O.transform_mat4([1.0, 2.0, -3.0], [2.0, 0.5, 1.5], [0.3, -1.1, 0.7]),
                       O.transform_mat4([0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.0, 0.0, 0.0])])
    tris["modelIndex"] = np.arange(len(tris)) % 2
    sid = ctx.upload(tris, None, s["materials"], models=models, flags=native.BUILD_KEEP_BINARY)
    wt, _ = O.model_to_world(tris, np.zeros(0, O.SPHERE), models)
    onodes, _, oslots, _, _ = O.build_lbvh(wt)
    assert _nodes_equal(ctx.nodes(sid), onodes)
    ctx.release(sid)


def test_supplied_node_buffer(ctx, native, O):
    """rt_scene_upload with the reference builder output handed in (no device build)."""
    tris, mats = small_cluster_scene(8, 1000)
    onodes, _, oslots, _, _ = O.build_lbvh(tris, flags=O.BUILD_REFERENCE_LEAVES)
    sid = ctx.upload(tris, None, mats, nodes=onodes.view(L.BVH_NODE), leaf_size=4)
    assert np.array_equal(ctx.slots(sid), oslots)
    ctx.release(sid)


def test_supplied_node_buffer_rejects_cycles(ctx, native, O):
    tris, mats = small_cluster_scene(1, 16)
    onodes, _, _, _, _ = O.build_lbvh(tris)
    bad = onodes.copy()
    bad[1]["left"] = 0  # points back at the root
    with pytest.raises(native.RTError) as e:
        ctx.upload(tris, None, mats, nodes=bad.view(L.BVH_NODE))
    assert e.value.status == 1
    bad = onodes.copy()
    bad[0]["right"] = 10 ** 6  # out of range
    with pytest.raises(native.RTError):
        ctx.upload(tris, None, mats, nodes=bad.view(L.BVH_NODE))


def test_empty_scene_and_bad_indices(ctx, native):
    sid = ctx.upload(None, None, W.make_materials([1, 1, 1], L.DIFFUSE))
    assert ctx.info(sid).nprim == 0
    ctx.release(sid)
    tris, mats = small_cluster_scene(1, 8)
    tris["materialIndex"][3] = 5
    with pytest.raises(native.RTError):
        ctx.upload(tris, None, mats)
    with pytest.raises(native.RTError):
        ctx.release(123456)


@pytest.mark.parametrize("width,leaf", [(2, 1), (4, 1), (4, 4), (8, 1), (8, 2)])
def test_wide_bvh_covers_every_slot_once(ctx, native, width, leaf):
    tris, mats = small_cluster_scene(8, 3000)
    sid = ctx.upload(tris, None, mats, width=width, leaf_size=leaf)
    inf = ctx.info(sid)
    wide = ctx.wide(sid).reshape(inf.nwide, 8 * width)
    child = wide[:, 6 * width:7 * width]
    seen = np.zeros(len(tris), np.int32)
    inner = 0
    for e in child.ravel():
        if e == 0xFFFFFFFF:
            continue
        if e & 0x80000000:
            first, cnt = e & 0x07FFFFFF, ((e >> 27) & 15) + 1
            assert cnt <= leaf
            seen[first:first + cnt] += 1
        else:
            inner += 1
            assert 0 < e < inf.nwide
    assert inner == inf.nwide - 1, "every wide node except the root has exactly one parent"
    assert np.all(seen == 1), "every leaf slot is referenced exactly once"
    assert inf.stack_bound == (width - 1) * inf.wide_depth + 1
    ctx.release(sid)
