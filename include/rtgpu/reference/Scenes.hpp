// Scenes.hpp -- stands in for the reference's Scenes.hpp (RaytracerGPU_MastersProject/Scenes.hpp:1-17):
// the five scene functions over this library's RaytraceScene.  With this directory on the
// quote-include path (g++ -iquote) and include/ on the include path, the reference's Scenes.cpp
// compiles unchanged against the library, provided the compiler does not find the reference's own
// headers first: quote includes search the including file's directory before -iquote, so the file
// is compiled from a directory without them (for example read from standard input: -x c++ -).
#pragma once

#include "rtgpu/glm.hpp"
#include "rtgpu/Scene.hpp"
