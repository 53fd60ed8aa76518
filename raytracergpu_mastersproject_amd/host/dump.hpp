// dump.hpp -- write a flattened RaytraceScene (model-space triangles, spheres, materials, model
// matrices) and render state as raw little-endian files, for inspection and parity tests.
#pragma once

#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "rtgpu/Scene.hpp"

namespace rtgpu_dump {

template <class T> inline void write_raw(const std::string &path, const std::vector<T> &v) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot write " + path);
    if (!v.empty()) f.write(reinterpret_cast<const char *>(v.data()), (std::streamsize)(sizeof(T) * v.size()));
}

inline void write_scene(const std::string &dir, const RaytraceScene &s) {
    write_raw(dir + "/tris.bin", s.getTriangles());
    write_raw(dir + "/spheres.bin", s.getSpheres());
    write_raw(dir + "/materials.bin", s.getMaterials());
    write_raw(dir + "/models.bin", s.getModels());
}

}  // namespace rtgpu_dump
