"""CPU-side checks of the boundary: the C-ABI library loads and exports every entry point declared in
include/rtgpu.h, the numpy layouts match the C header byte for byte (checked by compiling the header),
and host logic that needs no GPU behaves (errors, generators)."""
import re
import subprocess
import tempfile
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "rtgpu.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?(rt_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from raytracergpu_mastersproject_amd import native as N
    names = declared_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", str(N.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (rt_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    for n in names:
        assert hasattr(N.lib, n)
    assert set(N.EXPORTED) <= set(names)


def test_layouts_match_the_c_header():
    from raytracergpu_mastersproject_amd import layouts as L
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "rtgpu.h"
#define P(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
#define S(T) printf(#T " %zu\n", sizeof(T))
int main(void) {
  S(rt_triangle); P(rt_triangle, v1); P(rt_triangle, v2); P(rt_triangle, materialIndex); P(rt_triangle, modelIndex);
  S(rt_sphere); P(rt_sphere, radius); P(rt_sphere, materialIndex); P(rt_sphere, modelIndex);
  S(rt_material); P(rt_material, materialType);
  S(rt_bvh_node); P(rt_bvh_node, maxX); P(rt_bvh_node, left); P(rt_bvh_node, right); P(rt_bvh_node, primitiveIndex);
  P(rt_bvh_node, primitiveType);
  S(rt_params); P(rt_params, camLookAt); P(rt_params, verticalFOV); P(rt_params, numTriangles);
  P(rt_params, maxRayTraceDepth); P(rt_params, randomState);
  S(rt_ray); P(rt_ray, tmin); P(rt_ray, d); P(rt_ray, tmax); S(rt_hit); P(rt_hit, t);
  return 0;
}'''
    with tempfile.TemporaryDirectory() as d:
        src, exe = Path(d) / "l.c", Path(d) / "l"
        src.write_text(prog)
        subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
        got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True,
                                                                    text=True).stdout.split("\n") if line)
    dtypes = {"rt_triangle": L.TRIANGLE, "rt_sphere": L.SPHERE, "rt_material": L.MATERIAL, "rt_bvh_node": L.BVH_NODE,
              "rt_params": L.PARAMS, "rt_ray": L.RAY, "rt_hit": L.HIT}
    for key, val in got.items():
        if "." in key:
            t, f = key.split(".")
            assert dtypes[t].fields[f][1] == int(val), key
        else:
            assert dtypes[key].itemsize == int(val), key
    # the reference's sizes (definitions.glsl / SceneTypes.hpp / RaytracerBVH.hpp:31-42)
    assert got["rt_triangle"] == "64" and got["rt_sphere"] == "32" and got["rt_material"] == "32"
    assert got["rt_bvh_node"] == "40" and got["rt_params"] == "80"


def test_oracle_layouts_match_product_layouts():
    import oracle as O
    from raytracergpu_mastersproject_amd import layouts as L
    for a, b in [(O.TRIANGLE, L.TRIANGLE), (O.SPHERE, L.SPHERE), (O.MATERIAL, L.MATERIAL), (O.NODE, L.BVH_NODE),
                 (O.PARAMS, L.PARAMS), (O.RAY, L.RAY), (O.HIT, L.HIT)]:
        assert a.itemsize == b.itemsize and a.names == b.names


def test_errors_without_a_device_or_with_bad_arguments():
    from raytracergpu_mastersproject_amd import native as N
    import ctypes as C
    if N.device_count() == 0:
        h = C.c_void_p()
        assert N.lib.rt_create(C.byref(h), 0) != 0
        with pytest.raises(N.RTError):
            N.Context(0)
    assert N.lib.rt_create(None, 0) == 1  # RT_ERR_INVALID
    assert N.lib.rt_last_error(None) == b"null context"
    assert N.lib.rt_trace(None, 1, None, 0, None, None, 0, None, None) == 1


def test_workload_generators_are_deterministic():
    from raytracergpu_mastersproject_amd import layouts as L
    from raytracergpu_mastersproject_amd import workloads as W
    s1, s2 = W.soup_scene(), W.soup_scene()
    assert len(s1["tris"]) == 1028 and np.array_equal(s1["tris"], s2["tris"])
    a = W.SplitMix64(7).u64(100)
    b = np.zeros(100, np.uint64)
    W._genlib().rtw_u64(7, 0, 100, b.ctypes.data, 3)
    assert np.array_equal(a, b)
    t1, _ = W.cluster_scene(4, 1000, seed=9)
    t2, _ = W.cluster_scene(4, 1000, seed=9)
    assert np.array_equal(t1, t2) and len(t1) == 4000
    r = W.incoherent_rays(5000, [-1, -1, -1], [1, 1, 1], seed=2)
    assert np.allclose(np.linalg.norm(r["d"].astype(np.float64), axis=1), 1.0, atol=1e-6)
    assert r["o"].min() >= -1 and r["o"].max() <= 1 and np.all(r["tmin"] == np.float32(0.001))
    ico = W.icosphere_scene(freq=6)
    assert len(ico["tris"]) == 20 * 36 + 2
    mt = ico["materials"]["materialType"][:-1]
    assert set(np.unique(mt)) <= {L.DIFFUSE, L.METALLIC, L.DIELECTRIC}


def test_workload_sizes_of_the_headline_configs():
    from raytracergpu_mastersproject_amd import workloads as W
    assert 20 * 224 ** 2 == 1_003_520
    assert 64 * 65536 == 4_194_304


def test_ctypes_structs_match_the_c_header():
    """Every struct the Python binding passes across the ABI has the header's size and field offsets."""
    import ctypes as C
    from raytracergpu_mastersproject_amd import native as N
    structs = {"rt_build_options": N.BuildOptions, "rt_launch_config": N.LaunchConfig, "rt_scene_info": N.SceneInfo,
               "rt_batch_item": N.BatchItem, "rt_tune_result": N.TuneResult, "rt_render_job": N.RenderJob,
               "rt_server_config": N.ServerConfig, "rt_request": N.Request, "rt_request_stats": N.RequestStats,
               "rt_server_stats": N.ServerStats, "rt_tune_scene": N.TuneScene, "rt_build_candidate": N.BuildCandidate}
    lines = []
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    prog = "#include <stdio.h>\n#include <stddef.h>\n#include \"rtgpu.h\"\nint main(void) {\n" + "\n".join(lines) + \
        "\nreturn 0;\n}\n"
    with tempfile.TemporaryDirectory() as d:
        src, exe = Path(d) / "s.c", Path(d) / "s"
        src.write_text(prog)
        subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
        got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True,
                                                                    text=True).stdout.split("\n") if line)
    for cname, py in structs.items():
        assert C.sizeof(py) == int(got[cname]), cname
        for f, _ in py._fields_:
            assert getattr(py, f).offset == int(got[f"{cname}.{f}"]), f"{cname}.{f}"


def test_server_entry_points_reject_null_handles():
    from raytracergpu_mastersproject_amd import native as N
    assert N.lib.rt_server_create(None, None, None) == 1
    assert N.lib.rt_server_submit(None, None, None) == 1
    assert N.lib.rt_server_wait(None, 1, 0, None) == 1
    assert N.lib.rt_server_drain(None) == 1
    assert N.lib.rt_server_destroy(None) == 1
    assert N.lib.rt_server_last_error(None) == b"null server"
    assert N.lib.rt_render_batch(None, None, 0, None, None, None) == 1


def test_flag_constants_match_the_c_header():
    """Every RT_BUILD_/RT_TRACE_/RT_RENDER_/RT_REQ_/RT_SERVER_ #define has the same value in the binding."""
    import re
    from raytracergpu_mastersproject_amd import native as N
    text = (ROOT / "include" / "rtgpu.h").read_text()
    found = 0
    for name, value in re.findall(r"^#define RT_((?:BUILD|TRACE|RENDER|REQ|SERVER)_\w+) (0x[0-9A-Fa-f]+|\d+)u", text, re.M):
        assert getattr(N, name) == int(value, 0), name
        found += 1
    assert found >= 15
