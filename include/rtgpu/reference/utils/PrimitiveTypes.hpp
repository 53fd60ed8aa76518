// PrimitiveTypes.hpp -- stands in for utils/PrimitiveTypes.hpp: the reference's fixed-width aliases.
#pragma once

#include <cstddef>
#include <cstdint>

#include "rtgpu/SceneTypes.hpp"  // u32, f32

using u8 = std::uint8_t;
using u16 = std::uint16_t;
using u64 = std::uint64_t;
using i8 = std::int8_t;
using i16 = std::int16_t;
using i32 = std::int32_t;
using i64 = std::int64_t;
using f64 = double;
