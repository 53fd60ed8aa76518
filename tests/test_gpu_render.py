"""Path tracing on the GPU vs the oracle's restatement of raytraceBVH.comp (bit-exact images).

The soup image is compared with the oracle tracing every segment by brute force.  The other
images are compared with the oracle replaying the GPU's own wide BVH (method WIDE), which is
brute force except for rays with a zero direction component lying exactly in a box-face plane:
the traversal takes the zero as 2^-100 with its sign and culls the box where that nudged ray moves
out of it (pinned in test_oracle_bvh.py), and an odd image width puts the centre column's rays in
the plane x = 0 of such scenes."""
import numpy as np
import pytest

from raytracergpu_mastersproject_amd import layouts as L
from raytracergpu_mastersproject_amd import workloads as W

from helpers import upload_both, with_spheres

pytestmark = pytest.mark.gpu


def gpu_render(native, ctx, sid, params, w, h, spp, flags=0, accum=None, cfg=None):
    npix = w * h
    acc = native.DeviceArray(16 * npix)
    if accum is None:
        ctx.accum_clear(acc, npix)
    else:
        acc.upload(accum)
    segs = native.DeviceArray(8)
    segs.zero()
    ctx.render(sid, params, w, h, spp, acc, flags=flags, segments=segs, cfg=cfg)
    out = acc.to_numpy(np.float32, 4 * npix).reshape(npix, 4)
    return out, int(segs.to_numpy(np.uint64, 1)[0])


def test_soup_image_bit_exact(ctx, native, O):
    """The soup scene (1,024 triangles + ground + light), 256x256, 1 spp, depth 8: every pixel's
    accumulated RGBA (including the RNG carry in alpha) equals the oracle's bit for bit."""
    s = W.soup_scene()
    sid, osc = upload_both(ctx, O, s["tris"], None, s["materials"])
    g, gsegs = gpu_render(native, ctx, sid, s["params"], 256, 256, 1)
    o, osegs = osc.render(s["params"], 256, 256, 1)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    assert gsegs == osegs  # spp=1: no cached primaries
    rmse = float(np.sqrt(np.mean((g[:, :3] - o[:, :3]) ** 2)))
    assert rmse < 1e-4
    assert np.mean(g[:, 0] > 0) > 0.01
    ctx.release(sid)


@pytest.mark.parametrize("cache", [True, False])
def test_spp_chain_and_primary_cache(ctx, native, O, cache):
    """One lane per pixel (RENDER_NO_SPLIT): the cached first segment saves spp-1 traces per pixel."""
    s = W.soup_scene(width=64, height=48)
    sid, osc = upload_both(ctx, O, s["tris"], None, s["materials"])
    flags = (0 if cache else native.RENDER_NO_PRIMARY_CACHE) | native.RENDER_NO_SPLIT
    g4, segs4 = gpu_render(native, ctx, sid, s["params"], 64, 48, 4, flags)
    o4, osegs = osc.render(s["params"], 64, 48, 4, method=O.WIDE)
    assert np.array_equal(g4.view(np.uint32), o4.view(np.uint32))
    assert segs4 == (osegs - 64 * 48 * 3 if cache else osegs)
    # two successive calls of 2 spp continue the per-pixel chain exactly like one call of 4
    g2, _ = gpu_render(native, ctx, sid, s["params"], 64, 48, 2, flags)
    g22, _ = gpu_render(native, ctx, sid, s["params"], 64, 48, 2, flags, accum=g2)
    assert np.array_equal(g22.view(np.uint32), g4.view(np.uint32))
    ctx.release(sid)


@pytest.mark.parametrize("w,h,spp", [(64, 48, 5), (13, 29, 17), (8, 8, 64)])
def test_split_samples_equal_the_pixel_chain(ctx, native, O, w, h, spp):
    """One lane per (pixel, sample), summed in order afterwards, gives the oracle's image bit for bit
    (including the alpha carry), whether the image starts cleared or holds earlier samples; every
    sample traces its own first segment."""
    s = W.icosphere_scene(freq=10, width=w, height=h, spp=spp, depth=6)
    sid, osc = upload_both(ctx, O, s["tris"], None, s["materials"])
    g, segs = gpu_render(native, ctx, sid, s["params"], w, h, spp, native.RENDER_SPLIT_SAMPLES)
    o, osegs = osc.render(s["params"], w, h, spp, method=O.WIDE)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    assert segs == osegs
    # continuing an image: pixel-chain launch then split launch equals the oracle's 2*spp image
    g1, _ = gpu_render(native, ctx, sid, s["params"], w, h, spp, native.RENDER_NO_SPLIT)
    g2, _ = gpu_render(native, ctx, sid, s["params"], w, h, spp, native.RENDER_SPLIT_SAMPLES, accum=g1)
    o2, _ = osc.render(s["params"], w, h, 2 * spp, method=O.WIDE)
    assert np.array_equal(g2.view(np.uint32), o2.view(np.uint32))
    # the default picks the split layout for an image this small, with the same image
    gd, _ = gpu_render(native, ctx, sid, s["params"], w, h, spp)
    assert np.array_equal(gd.view(np.uint32), o.view(np.uint32))
    ctx.release(sid)


def _material_scene():
    s = W.icosphere_scene(freq=12, width=96, height=64, spp=2, depth=6)
    return s


@pytest.mark.parametrize("ref_mats", [False, True])
def test_metal_dielectric_scene(ctx, native, O, ref_mats):
    s = _material_scene()
    sid, osc = upload_both(ctx, O, s["tris"], None, s["materials"])
    flags = native.RENDER_REFERENCE_MATERIALS if ref_mats else 0
    g, _ = gpu_render(native, ctx, sid, s["params"], 96, 64, 2, flags)
    o, _ = osc.render(s["params"], 96, 64, 2, flags=O.RENDER_REFERENCE_MATERIALS if ref_mats else 0, method=O.WIDE)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    ctx.release(sid)


def test_spheres_scene(ctx, native, O):
    s = W.soup_scene(ntri=300, width=80, height=60)
    sph = with_spheres(s["tris"], s["materials"], 6)
    sid, osc = upload_both(ctx, O, s["tris"], sph, s["materials"])
    g, _ = gpu_render(native, ctx, sid, s["params"], 80, 60, 2)
    o, _ = osc.render(s["params"], 80, 60, 2, method=O.WIDE)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    ctx.release(sid)


def test_rows_beyond_600_and_odd_sizes(ctx, native, O):
    """Seed aliasing at rows >= 600 (random.glsl:10 uses a literal 600) and non-multiple-of-8 sizes."""
    s = W.soup_scene(ntri=128)
    sid, osc = upload_both(ctx, O, s["tris"], None, s["materials"])
    w, h = 13, 611
    g, _ = gpu_render(native, ctx, sid, s["params"], w, h, 1)
    o, _ = osc.render(s["params"], w, h, 1, method=O.WIDE)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    ctx.release(sid)


def test_resolve(ctx, native, O):
    rng = np.random.default_rng(0)
    acc = rng.uniform(-0.5, 40.0, (1000, 4)).astype(np.float32)
    da = native.DeviceArray.from_numpy(acc)
    do = native.DeviceArray(acc.nbytes)
    ctx.resolve(da, 1000, 7, do)
    g = do.to_numpy(np.float32, 4000).reshape(1000, 4)
    o = O.resolve(acc, 7)
    assert np.array_equal(np.nan_to_num(g, nan=-1.0).view(np.uint32), np.nan_to_num(o, nan=-1.0).view(np.uint32))


def test_depth_zero_and_empty_scene(ctx, native, O):
    s = W.soup_scene(ntri=64, width=16, height=16)
    p = s["params"].copy()
    p["maxRayTraceDepth"] = 0
    sid, osc = upload_both(ctx, O, s["tris"], None, s["materials"])
    g, segs = gpu_render(native, ctx, sid, p, 16, 16, 3)
    o, _ = osc.render(p, 16, 16, 3, method=O.WIDE)
    assert segs == 0 and np.array_equal(g.view(np.uint32), o.view(np.uint32))
    ctx.release(sid)
    sid = ctx.upload(None, None, s["materials"])
    g, _ = gpu_render(native, ctx, sid, s["params"], 16, 16, 2)
    assert np.all(g[:, :3] == 0.0)
    ctx.release(sid)


@pytest.mark.parametrize("cache", [True, False])
def test_wavefront_equals_the_oracle(ctx, native, O, cache, monkeypatch):
    """Bounce-by-bounce rendering (ray-query launches over a queue of segments, then shading)
    gives the oracle's image bit for bit, in one pass or in passes of 2 samples that continue
    the carry chain; with the primary cache a pixel's primary segment is traced once per pass."""
    w, h, spp = 64, 48, 5
    s = W.icosphere_scene(freq=10, width=w, height=h, spp=spp, depth=6)
    sid, osc = upload_both(ctx, O, s["tris"], None, s["materials"])
    flags = native.RENDER_WAVEFRONT | (0 if cache else native.RENDER_NO_PRIMARY_CACHE)
    o, osegs = osc.render(s["params"], w, h, spp, method=O.WIDE)
    g, segs = gpu_render(native, ctx, sid, s["params"], w, h, spp, flags)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    assert segs == (osegs - w * h * (spp - 1) if cache else osegs)
    # the late bounces run in one persistent launch from bounce 1 by default (bounce 0 is the
    # shading pass over the primary hits); from bounce 2 or 3 (the bounces before it shaded by
    # their own launches), or never (every bounce its own launches), the image and the traced
    # segments are the same
    for finish in ("2", "3", "99"):
        monkeypatch.setenv("RTGPU_WAVE_FINISH", finish)
        gf, segf = gpu_render(native, ctx, sid, s["params"], w, h, spp, flags)
        assert np.array_equal(gf.view(np.uint32), o.view(np.uint32)), finish
        assert segf == segs, finish
    monkeypatch.delenv("RTGPU_WAVE_FINISH")
    monkeypatch.setenv("RTGPU_WAVE_PATHS", str(2 * w * h))  # passes of 2, 2 and 1 samples
    g3, segs3 = gpu_render(native, ctx, sid, s["params"], w, h, spp, flags)
    assert np.array_equal(g3.view(np.uint32), o.view(np.uint32))
    assert segs3 == (osegs - w * h * (spp - 3) if cache else osegs)
    # continuing an image: a persistent launch of spp samples, then spp bounce-by-bounce samples
    g1, _ = gpu_render(native, ctx, sid, s["params"], w, h, spp, native.RENDER_PERSISTENT_PATHS)
    g2, _ = gpu_render(native, ctx, sid, s["params"], w, h, spp, flags, accum=g1)
    o2, _ = osc.render(s["params"], w, h, 2 * spp, method=O.WIDE)
    assert np.array_equal(g2.view(np.uint32), o2.view(np.uint32))
    ctx.release(sid)


def test_wavefront_materials_spheres_and_edges(ctx, native, O):
    """The bounce-by-bounce schedule on metal/dielectric materials (both material modes), sphere
    leaves, rows past the seed aliasing at 600, and depth 0 (nothing traced)."""
    wf = native.RENDER_WAVEFRONT
    s = _material_scene()
    sid, osc = upload_both(ctx, O, s["tris"], None, s["materials"])
    for ref in (False, True):
        g, _ = gpu_render(native, ctx, sid, s["params"], 96, 64, 2,
                          wf | (native.RENDER_REFERENCE_MATERIALS if ref else 0))
        o, _ = osc.render(s["params"], 96, 64, 2, flags=O.RENDER_REFERENCE_MATERIALS if ref else 0, method=O.WIDE)
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), ref
    ctx.release(sid)
    s = W.soup_scene(ntri=300, width=80, height=60)
    sph = with_spheres(s["tris"], s["materials"], 6)
    sid, osc = upload_both(ctx, O, s["tris"], sph, s["materials"])
    g, _ = gpu_render(native, ctx, sid, s["params"], 80, 60, 2, wf)
    o, _ = osc.render(s["params"], 80, 60, 2, method=O.WIDE)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    ctx.release(sid)
    s = W.soup_scene(ntri=128)
    sid, osc = upload_both(ctx, O, s["tris"], None, s["materials"])
    g, _ = gpu_render(native, ctx, sid, s["params"], 13, 611, 1, wf)
    o, _ = osc.render(s["params"], 13, 611, 1, method=O.WIDE)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    p = s["params"].copy()
    p["maxRayTraceDepth"] = 0
    g, segs = gpu_render(native, ctx, sid, p, 16, 16, 3, wf)
    o, _ = osc.render(p, 16, 16, 3, method=O.WIDE)
    assert segs == 0 and np.array_equal(g.view(np.uint32), o.view(np.uint32))
    with pytest.raises(native.RTError):
        gpu_render(native, ctx, sid, s["params"], 16, 16, 1, wf | native.RENDER_NO_SPLIT)
    ctx.release(sid)


def test_default_schedule_of_a_large_image(ctx, native, O):
    """An image of at least 2^22 pixel-samples is rendered bounce by bounce by default, with the
    same image as the persistent launch (which the tests above pin to the oracle)."""
    s = W.icosphere_scene(freq=16, width=512, height=512, spp=16, depth=8)
    sid = ctx.upload(s["tris"], None, s["materials"])
    gd, segd = gpu_render(native, ctx, sid, s["params"], 512, 512, 16)
    gp, segp = gpu_render(native, ctx, sid, s["params"], 512, 512, 16, native.RENDER_PERSISTENT_PATHS)
    assert np.array_equal(gd.view(np.uint32), gp.view(np.uint32))
    assert 0 < segd < segp  # primaries once per pixel instead of once per work item
    ctx.release(sid)


@pytest.mark.parametrize("width", [2, 8])
def test_wavefront_other_bvh_widths_and_an_empty_scene(ctx, native, O, width):
    """Bounce-by-bounce rendering over BVH4's siblings (W2, W8) equals the oracle; an empty scene
    renders black without tracing anything."""
    w, h, spp = 40, 30, 3
    s = W.icosphere_scene(freq=8, width=w, height=h, spp=spp, depth=5)
    sid, osc = upload_both(ctx, O, s["tris"], None, s["materials"], width=width)
    g, _ = gpu_render(native, ctx, sid, s["params"], w, h, spp, native.RENDER_WAVEFRONT)
    o, _ = osc.render(s["params"], w, h, spp, method=O.WIDE)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))
    ctx.release(sid)
    sid = ctx.upload(None, None, s["materials"])
    g, segs = gpu_render(native, ctx, sid, s["params"], 16, 16, 2, native.RENDER_WAVEFRONT)
    assert np.all(g[:, :3] == 0.0) and segs == 16 * 16  # primaries traced (and missed), nothing after
    ctx.release(sid)


def test_renders_on_two_streams_share_the_scratch_cache(ctx, native, O):
    """Bounce-by-bounce renders enqueued on two streams at once take their path storage from the
    context's scratch cache (a block returns with an event; another stream reuses it only after
    that event); every image equals the same render done alone."""
    s = W.icosphere_scene(freq=10, width=96, height=72, spp=6, depth=6)
    sid = ctx.upload(s["tris"], None, s["materials"])
    w, h = 96, 72
    states = [3, 5, 8, 13, 21, 34]
    alone = []
    for rs in states:
        p = s["params"].copy()
        p["randomState"] = rs
        g, _ = gpu_render(native, ctx, sid, p, w, h, 6, native.RENDER_WAVEFRONT)
        alone.append(g)
    streams = [native.Stream(), native.Stream()]
    accs = []
    for k, rs in enumerate(states):  # alternate streams; blocks are reused within a stream at once
        p = s["params"].copy()
        p["randomState"] = rs
        acc = native.DeviceArray(16 * w * h)
        ctx.accum_clear(acc, w * h, stream=streams[k % 2])
        ctx.render(sid, p, w, h, 6, acc, flags=native.RENDER_WAVEFRONT, stream=streams[k % 2])
        accs.append(acc)
    for st in streams:
        st.synchronize()
    for acc, ref in zip(accs, alone):
        got = acc.to_numpy(np.float32, 4 * w * h).reshape(-1, 4)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    ctx.release(sid)
