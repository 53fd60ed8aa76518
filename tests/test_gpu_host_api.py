"""The C++ render entry (RaytracerBVHRenderer::Raytracer, headless) on the GPU: one frame of a reference
scene through Scenes -> RaytraceScene -> rt_scene_upload (device model->world + LBVH) -> rt_render, compared
bit for bit with the oracle rendering the same flattened scene."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from raytracergpu_mastersproject_amd import layouts as L

pytestmark = pytest.mark.gpu

PKG = Path(__file__).resolve().parent.parent / "raytracergpu_mastersproject_amd"


@pytest.mark.parametrize("scene,size,spp,depth", [("cornell", (48, 40), 2, 6), ("simple", (40, 32), 3, 8),
                                                   ("mixed", (32, 32), 1, 12), ("spheres", (40, 24), 2, 5)])
def test_cpp_raytracer_frame_matches_oracle(native, tmp_path, scene, size, spp, depth):
    w, h = size
    out = subprocess.run([str(PKG / "rtgpu_render"), "--scene", scene, "--size", f"{w}x{h}", "--spp", str(spp),
                          "--depth", str(depth), "--seed", "5", "--frames", "2", "--quiet", "--out",
                          str(tmp_path / "img.ppm"), "--dump", str(tmp_path)], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    meta = json.loads(out.stdout.strip().splitlines()[-1])
    tris = np.fromfile(tmp_path / "tris.bin", L.TRIANGLE)
    sph = np.fromfile(tmp_path / "spheres.bin", L.SPHERE)
    mats = np.fromfile(tmp_path / "materials.bin", L.MATERIAL)
    models = np.fromfile(tmp_path / "models.bin", np.float32).reshape(-1, 16)
    params = np.fromfile(tmp_path / "params.bin", L.PARAMS)[0]
    accum = np.fromfile(tmp_path / "accum.bin", np.float32).reshape(-1, 4)
    assert meta["triangles"] == len(tris) and params["maxRayTraceDepth"] == depth
    wt, ws = O.model_to_world(tris, sph, models)
    osc = O.OracleScene(wt, ws, mats)
    oacc, _ = osc.render(params, w, h, spp)
    assert np.array_equal(accum.view(np.uint32), oacc.view(np.uint32))
    ppm = (tmp_path / "img.ppm").read_bytes()
    assert ppm.startswith(f"P6\n{w} {h}\n255\n".encode()) and len(ppm) == len(f"P6\n{w} {h}\n255\n") + 3 * w * h


def test_spp_sweep_writes_runtimes_csv(native, tmp_path):
    """Config::RunRayPerPixelIncreasingDemo (Config.hpp:17-23, RaytracerBVH.hpp:522-570): rays per pixel from 1,
    +2 every 2 frames while at most 5, one mean frame time per level in runtimes.csv."""
    csv = tmp_path / "runtimes.csv"
    out = subprocess.run([str(PKG / "rtgpu_render"), "--scene", "simple", "--size", "32x32", "--quiet",
                          "--spp-sweep", "2,1,5,2", "--csv", str(csv), "--out", str(tmp_path / "img.ppm")],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    meta = json.loads(out.stdout.strip().splitlines()[0])
    lines = csv.read_text().splitlines()
    assert meta["spp_sweep_levels"] == 3 and len(lines) == 3  # 1, 3 and 5 rays per pixel
    for k, ln in enumerate(lines):
        idx, t, rest = ln.split(",")
        assert int(idx) == k + 1 and t.strip().endswith("us") and int(t.strip()[:-2]) > 0 and rest == ""
    assert json.loads(out.stdout.strip().splitlines()[-1])["spp"] == 7  # the level past the maximum
