// types.hpp -- scalar and enum types of the wake C++ API.
//
// The names and enumerator values match include/wake/core/types.hpp:12-99 of
// the reference (real_t is float; CellType codes are the uint8 values the
// kernels read), so code written against the reference compiles unchanged.
#pragma once

#include <cstddef>
#include <cstdint>

namespace wake {

using real_t = float;
using index_t = std::size_t;
using int_coord_t = int32_t;

struct Vec2 {
  real_t x{0.0f}, y{0.0f};
  constexpr Vec2() = default;
  constexpr Vec2(real_t a, real_t b) : x(a), y(b) {}
  constexpr Vec2 operator+(const Vec2& o) const { return {x + o.x, y + o.y}; }
  constexpr Vec2 operator-(const Vec2& o) const { return {x - o.x, y - o.y}; }
  constexpr Vec2 operator*(real_t s) const { return {x * s, y * s}; }
  constexpr Vec2 operator/(real_t s) const { return {x / s, y / s}; }
  constexpr real_t dot(const Vec2& o) const { return x * o.x + y * o.y; }
  constexpr real_t norm_sq() const { return x * x + y * y; }
};

struct Vec3 {
  real_t x{0.0f}, y{0.0f}, z{0.0f};
  constexpr Vec3() = default;
  constexpr Vec3(real_t a, real_t b, real_t c) : x(a), y(b), z(c) {}
  constexpr Vec3 operator+(const Vec3& o) const { return {x + o.x, y + o.y, z + o.z}; }
  constexpr Vec3 operator-(const Vec3& o) const { return {x - o.x, y - o.y, z - o.z}; }
  constexpr Vec3 operator*(real_t s) const { return {x * s, y * s, z * s}; }
  constexpr Vec3 operator/(real_t s) const { return {x / s, y / s, z / s}; }
  Vec3& operator+=(const Vec3& o) {
    x += o.x;
    y += o.y;
    z += o.z;
    return *this;
  }
  Vec3& operator-=(const Vec3& o) {
    x -= o.x;
    y -= o.y;
    z -= o.z;
    return *this;
  }
  Vec3& operator*=(real_t s) {
    x *= s;
    y *= s;
    z *= s;
    return *this;
  }
  constexpr real_t dot(const Vec3& o) const { return x * o.x + y * o.y + z * o.z; }
  constexpr real_t norm_sq() const { return x * x + y * y + z * z; }
  constexpr Vec3 cross(const Vec3& o) const { return {y * o.z - z * o.y, z * o.x - x * o.z, x * o.y - y * o.x}; }
};

enum class CellType : uint8_t { Fluid = 0, Solid = 1, Inflow = 2, Outflow = 3, Periodic = 4, Symmetry = 5, MovingWall = 6 };

enum class CollisionType : uint8_t { BGK = 0, TRT = 1, MRT = 2, Entropic = 3 };

// Population storage in HBM: FP32, or FP16 / BF16 of f_q - w_q (FP32 arithmetic).
enum class PopulationStorage : uint8_t { F32 = 0, F16 = 1, BF16 = 2 };

}  // namespace wake
