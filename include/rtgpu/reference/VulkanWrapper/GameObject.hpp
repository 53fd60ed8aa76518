// GameObject.hpp -- stands in for VulkanWrapper/GameObject.hpp (GameObject, TransformComponent;
// GameObject.hpp:1-301), which the reference's Scenes.cpp includes (Scenes.cpp:5).
#pragma once

#include "rtgpu/glm.hpp"
#include "rtgpu/Scene.hpp"
