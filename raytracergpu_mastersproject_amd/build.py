"""Build recipe for the native library (librtgpu.so) and the host C++ tools.

Compiles every HIP source for gfx950 with hipcc, in-tree, so the built shared
library travels with the source tree.  Usage: ``python -m raytracergpu_mastersproject_amd.build``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"
LIB = PKG / "librtgpu.so"
GEN_SRC = HOST / "workloads_gen.cpp"
GEN_LIB = PKG / "libworkloads.so"
TOOLS = {"rtgpu_render": HOST / "main.cpp", "rtgpu_scene_dump": HOST / "scene_dump.cpp"}

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RTGPU_ARCH", "gfx950")

# -ffp-contract=off: no FMA contraction, so every fp32 op rounds once (bit parity with the CPU oracle).
# -packed-fp32-ops is disabled for device code (packed fp32 forms are avoided on this hardware).
COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{CSRC}", f"-I{HOST}",
]
HOST_CXX = ["g++", "-O2", "-fPIC", "-std=c++20", "-ffp-contract=off", "-Wall", f"-I{INCLUDE}", f"-I{HOST}"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}")
    noisy = [ln for ln in proc.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if noisy:
        sys.stderr.write("\n".join(noisy) + "\n")


def sources() -> tuple[list[Path], list[Path]]:
    hip = sorted(CSRC.glob("*.hip"))
    tool_srcs = {p.name for p in TOOLS.values()}
    cpp = sorted(p for p in HOST.glob("*.cpp") if p.name not in tool_srcs and p.name != GEN_SRC.name)
    return hip, cpp


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    """Compile all sources and link librtgpu.so (always recompiles)."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    BUILD.mkdir(exist_ok=True)
    hip, cpp = sources()
    jobs = jobs or min(16, max(1, os.cpu_count() or 1))
    cmds = []
    objs = []
    for src in hip:
        obj = BUILD / (src.stem + ".o")
        cmds.append([HIPCC, *COMMON, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    for src in cpp:
        obj = BUILD / ("host_" + src.stem + ".o")
        cmds.append([*HOST_CXX, "-c", str(src), "-o", str(obj)])
        objs.append(obj)
    gen_tmp = GEN_LIB.with_suffix(".so.tmp")
    cmds.append([*HOST_CXX, "-O3", "-shared", str(GEN_SRC), "-o", str(gen_tmp), "-pthread"])
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for fut in [ex.submit(_run, c) for c in cmds]:
            fut.result()
    os.replace(gen_tmp, GEN_LIB)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    for name, main in TOOLS.items():
        if main.exists():
            _run([*HOST_CXX, str(main), "-o", str(PKG / name), f"-L{PKG}", "-lrtgpu",
                  f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"])
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
