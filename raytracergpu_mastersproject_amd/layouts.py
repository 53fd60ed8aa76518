"""numpy dtypes for the byte layouts of include/rtgpu.h.

Each mirrors one of the reference's GPU structs (RaytracerGPU_MastersProject/
shaders/include/definitions.glsl, VulkanWrapper/SceneTypes.hpp) byte for byte.
"""
import numpy as np

TRIANGLE = np.dtype([("v0", "<f4", 4), ("v1", "<f4", 4), ("v2", "<f4", 4),
                     ("materialIndex", "<u4"), ("modelIndex", "<u4"), ("_pad", "<u4", 2)])
SPHERE = np.dtype([("center", "<f4", 4), ("radius", "<f4"), ("materialIndex", "<u4"),
                   ("modelIndex", "<u4"), ("_pad", "<u4")])
MATERIAL = np.dtype([("albedo", "<f4", 4), ("materialType", "<u4"), ("_pad", "<u4", 3)])
BVH_NODE = np.dtype([("minX", "<f4"), ("maxX", "<f4"), ("minY", "<f4"), ("maxY", "<f4"),
                     ("minZ", "<f4"), ("maxZ", "<f4"), ("left", "<u4"), ("right", "<u4"),
                     ("primitiveIndex", "<u4"), ("primitiveType", "<u4")])
PARAMS = np.dtype([("camPos", "<f4", 4), ("camLookAt", "<f4", 4), ("camUpDir", "<f4", 4),
                   ("verticalFOV", "<f4"), ("numTriangles", "<u4"), ("numSpheres", "<u4"),
                   ("numMaterials", "<u4"), ("numLights", "<u4"), ("maxRayTraceDepth", "<u4"),
                   ("randomState", "<u4"), ("_pad", "<u4")])
RAY = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT = np.dtype([("prim", "<u4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])

assert TRIANGLE.itemsize == 64 and SPHERE.itemsize == 32 and MATERIAL.itemsize == 32
assert BVH_NODE.itemsize == 40 and PARAMS.itemsize == 80 and RAY.itemsize == 32 and HIT.itemsize == 16

LIGHT, DIFFUSE, METALLIC, DIELECTRIC = 0, 1, 2, 3
SPHERE_PRIMITIVE, TRIANGLE_PRIMITIVE = 0, 1
NO_HIT = 0xFFFFFFFF
SPHERE_BIT = 0x80000000


def make_params(cam_pos, look_at, up, vfov, ntri, nsph, nmat, depth, random_state, num_lights=20):
    """Fill a ParameterUBO record (RaytracerBVH.hpp:340-352 fills the same fields)."""
    p = np.zeros((), dtype=PARAMS)
    p["camPos"][:3] = cam_pos
    p["camLookAt"][:3] = look_at
    p["camUpDir"][:3] = up
    p["verticalFOV"] = vfov
    p["numTriangles"] = ntri
    p["numSpheres"] = nsph
    p["numMaterials"] = nmat
    p["numLights"] = num_lights
    p["maxRayTraceDepth"] = depth
    p["randomState"] = random_state
    return p
