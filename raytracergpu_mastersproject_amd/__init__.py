"""MI355X-native BVH ray tracing: the hot path of RaytracerGPU_MastersProject on gfx950.

Python access to the C ABI (include/rtgpu.h) for tests and benchmarks; the
reference-shaped C++ API (RaytraceScene, GameObject, Scenes, Raytracer) lives in
host/.  Importing ``native`` loads librtgpu.so and raises if it is not built.
"""
from . import layouts  # noqa: F401

__all__ = ["layouts", "native", "workloads"]
