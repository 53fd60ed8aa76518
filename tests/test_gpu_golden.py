"""The GPU against the committed golden answer: the soup workload read from tests/golden/ (not regenerated),
rendered at 256x256, 1 spp, depth 8 through rt_scene_upload + rt_render, must reproduce the committed
accumulation digest bit for bit."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from raytracergpu_mastersproject_amd import layouts as L

pytestmark = pytest.mark.gpu
G = Path(__file__).resolve().parent / "golden"


def test_soup_render_matches_golden_digest(ctx, native):
    g = json.loads((G / "golden.json").read_text())["soup_image"]
    tris = np.fromfile(G / "soup_tris.bin", L.TRIANGLE)
    mats = np.fromfile(G / "soup_materials.bin", L.MATERIAL)
    params = np.fromfile(G / "soup_params.bin", L.PARAMS)[0]
    sid = ctx.upload(tris, None, mats)
    w, h = g["width"], g["height"]
    acc = native.DeviceArray(16 * w * h)
    segs = native.DeviceArray(8)
    segs.zero()
    ctx.accum_clear(acc, w * h)
    ctx.render(sid, params, w, h, g["spp"], acc, segments=segs)
    img = acc.to_numpy(np.float32, 4 * w * h).reshape(-1, 4)
    assert hashlib.sha256(img.tobytes()).hexdigest() == g["accum_sha256"]
    assert int(segs.to_numpy(np.uint64, 1)[0]) == g["segments"]
    ctx.release(sid)
