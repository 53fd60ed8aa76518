// glm.hpp -- the few glm names the reference's scene descriptions use (RaytracerGPU_MastersProject/
// Scenes.cpp: glm::vec3, glm::pi<T>(), glm::radians), so that file compiles against this library
// without glm.  The values are glm 0.9.9's: pi<T>() and radians() are the same constants rounded to
// T, and vec3 is the library's three-float vector, constructible from any three arithmetic values.
#pragma once

#include <type_traits>

#include "SceneTypes.hpp"

namespace glm {

using vec3 = rtgpu::vec3;

template <typename T> constexpr T pi() {
    static_assert(std::is_floating_point_v<T>, "glm::pi needs a floating-point type");
    return static_cast<T>(3.14159265358979323846264338327950288);
}

template <typename T> constexpr T radians(T degrees) {
    static_assert(std::is_floating_point_v<T>, "glm::radians needs a floating-point type");
    return degrees * static_cast<T>(0.01745329251994329576923690768489);
}

}  // namespace glm
