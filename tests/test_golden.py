"""The committed golden vectors (tests/golden/, written by tests/golden/make_golden.py) against the oracle
and the workload generators as they are now: RNG sequences of random.glsl, the soup workload's inputs and
its rendered image, and the digests of the large generated workloads."""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from raytracergpu_mastersproject_amd import layouts as L
from raytracergpu_mastersproject_amd import workloads as W

G = Path(__file__).resolve().parent / "golden"
GOLD = json.loads((G / "golden.json").read_text())


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("case", GOLD["rng"], ids=lambda c: f"{c['x']},{c['y']},{c['random_state']}")
def test_rng_sequences(case):
    seed = O.rng_seed(case["x"], case["y"], case["random_state"])
    assert seed == case["seed"] == (600 * case["x"] + case["y"]) * (case["random_state"] + 1) % 2**32
    draws, state = O.rng_sequence(seed, len(case["draws_f32_bits"]))
    assert [int(v) for v in draws.view(np.uint32)] == case["draws_f32_bits"]
    assert state == case["state_after"]


def test_soup_inputs_and_image():
    g = GOLD["soup_image"]
    s = W.soup_scene()
    tris = np.fromfile(G / "soup_tris.bin", L.TRIANGLE)
    mats = np.fromfile(G / "soup_materials.bin", L.MATERIAL)
    params = np.fromfile(G / "soup_params.bin", L.PARAMS)[0]
    assert np.array_equal(tris, s["tris"]) and np.array_equal(mats, s["materials"])
    assert params.tobytes() == np.ascontiguousarray(s["params"]).tobytes()
    assert g["inputs_sha256"] == {"tris": sha(tris), "materials": sha(mats), "params": sha(params)}
    acc, segs = O.OracleScene(tris, None, mats).render(params, g["width"], g["height"], g["spp"], method=O.REF_BVH,
                                                      threads=8)
    assert segs == g["segments"] and sha(acc) == g["accum_sha256"]
    for i, bits in g["pixels"].items():
        assert [int(v) for v in acc[int(i)].view(np.uint32)] == bits
    assert np.mean(acc[:, 0] > 0) > 0.01  # the emitter lights part of the image


def test_generator_digests():
    g = GOLD["generators"]
    tris, _ = W.cluster_scene(64, 65536, seed=2)
    assert sha(tris) == g["cluster_scene_64x65536_seed2"]
    lo, hi = W.scene_bounds(tris)
    assert [list(map(float, lo)), list(map(float, hi))] == g["scene_bounds"]
    assert sha(W.incoherent_rays(1 << 20, lo, hi, seed=3)) == g["incoherent_rays_first_1048576_seed3"]
    ico = W.icosphere_scene(freq=32)
    assert sha(ico["tris"]) == g["icosphere_freq32_tris"] and sha(ico["materials"]) == g["icosphere_freq32_materials"]
    assert [int(v) for v in W.serving_scene_sizes()] == g["serving_scene_sizes"]
    assert W.serving_requests(8) == g["serving_requests_8"]
