// SceneTypes.hpp -- host-side scene types with the reference's names and GPU layouts
// (RaytracerGPU_MastersProject/VulkanWrapper/SceneTypes.hpp:13-127), without glm.
#pragma once

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "../rtgpu.h"

using u32 = uint32_t;
using f32 = float;

namespace rtgpu {

struct vec3 {
    f32 x = 0.0f, y = 0.0f, z = 0.0f;
    constexpr vec3() = default;
    constexpr vec3(f32 a, f32 b, f32 c) : x(a), y(b), z(c) {}
    // any three arithmetic values, each converted to f32 (glm's vec3(X, Y, Z) constructor)
    template <typename A, typename B, typename C,
              typename = std::enable_if_t<std::is_arithmetic_v<A> && std::is_arithmetic_v<B> &&
                                          std::is_arithmetic_v<C>>>
    constexpr vec3(A a, B b, C c) : x(static_cast<f32>(a)), y(static_cast<f32>(b)), z(static_cast<f32>(c)) {}
    constexpr explicit vec3(f32 s) : x(s), y(s), z(s) {}
    constexpr f32 operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    bool operator==(const vec3 &o) const { return x == o.x && y == o.y && z == o.z; }
};

// column-major 4x4 matrix (glm convention): m[col][row]
struct mat4 {
    f32 m[4][4] = {};
    static mat4 identity() {
        mat4 r;
        for (int i = 0; i < 4; ++i) r.m[i][i] = 1.0f;
        return r;
    }
    const f32 *data() const { return &m[0][0]; }
};

struct mat3 {
    f32 m[3][3] = {};
};

constexpr f32 pi() { return 3.14159265358979323846f; }
inline f32 radians(f32 deg) { return deg * (pi() / 180.0f); }

}  // namespace rtgpu

namespace SceneTypes {

enum class MaterialType : u32 { LIGHT = RT_LIGHT, DIFFUSE = RT_DIFFUSE, METALLIC = RT_METALLIC, DIELECTRIC = RT_DIELECTRIC };

namespace CPU {
struct Triangle {
    rtgpu::vec3 v0, v1, v2;
};
struct Sphere {
    rtgpu::vec3 center;
    f32 radius;
};
}  // namespace CPU

namespace GPU {
// 64-byte matrix of one game object (definitions.glsl:6-8)
struct Model {
    rtgpu::mat4 modelMatrix;
};

using Triangle = rt_triangle;
using Sphere = rt_sphere;

// definitions.glsl:10-13; `extra` occupies albedo.w (fuzz for METALLIC, index of refraction for DIELECTRIC)
struct Material {
    rtgpu::vec3 albedo{0.0f};
    f32 extra = 0.0f;
    MaterialType materialType = MaterialType::DIFFUSE;
    u32 _pad[3] = {0, 0, 0};
    Material() = default;
    Material(rtgpu::vec3 color, MaterialType type, f32 fuzz_or_ior = 0.0f)
        : albedo(color), extra(fuzz_or_ior), materialType(type) {}
    bool operator==(const Material &o) const {
        return albedo == o.albedo && extra == o.extra && materialType == o.materialType;
    }
};
static_assert(sizeof(Material) == sizeof(rt_material), "Material layout");
static_assert(sizeof(Model) == 64, "Model layout");

using BVHNode = rt_bvh_node;
}  // namespace GPU

using Material = GPU::Material;

}  // namespace SceneTypes
