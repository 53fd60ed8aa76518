// RTModel.hpp -- stands in for VulkanWrapper/RTModel.hpp (RTModel, loadModel; RTModel.hpp:1-85),
// which the reference's Scenes.cpp includes (Scenes.cpp:4).
#pragma once

#include "rtgpu/glm.hpp"
#include "rtgpu/Scene.hpp"
