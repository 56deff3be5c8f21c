/*
 * Vector types for the fluid:: API. libfluid's headers use glm::tvec3 /
 * glm::tvec4 (include/fluid/fluid.hpp:15-16); when <glm/glm.hpp> is
 * available this header uses it, so user code keeps passing glm vectors.
 * Otherwise it provides a minimal glm-compatible subset in namespace glm
 * (members x,y,z[,w]; elementwise + - * /; distance, length, normalize,
 * clamp, to_string) with the same layout (3 or 4 packed T).
 */
#ifndef LIBFLUID_AMD_VEC_HPP
#define LIBFLUID_AMD_VEC_HPP

#if defined(__has_include)
#if __has_include(<glm/glm.hpp>) && !defined(LIBFLUID_AMD_NO_GLM)
#define LIBFLUID_AMD_HAS_GLM 1
#endif
#endif

#ifdef LIBFLUID_AMD_HAS_GLM
#ifndef GLM_ENABLE_EXPERIMENTAL
#define GLM_ENABLE_EXPERIMENTAL
#endif
#include <glm/glm.hpp>
#include <glm/gtx/normal.hpp>
#include <glm/gtx/string_cast.hpp>
#else

#include <algorithm>
#include <cmath>
#include <string>

namespace glm {

template <typename T> struct tvec3 {
  T x, y, z;
  tvec3() : x(T(0)), y(T(0)), z(T(0)) {}
  explicit tvec3(T s) : x(s), y(s), z(s) {}
  tvec3(T a, T b, T c) : x(a), y(b), z(c) {}
  template <typename U> explicit tvec3(const tvec3<U>& o) : x(T(o.x)), y(T(o.y)), z(T(o.z)) {}
  T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
  const T& operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
  tvec3& operator+=(const tvec3& o) { x += o.x; y += o.y; z += o.z; return *this; }
  tvec3& operator-=(const tvec3& o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
  bool operator==(const tvec3& o) const { return x == o.x && y == o.y && z == o.z; }
  bool operator!=(const tvec3& o) const { return !(*this == o); }
};

template <typename T> struct tvec4 {
  T x, y, z, w;
  tvec4() : x(T(0)), y(T(0)), z(T(0)), w(T(0)) {}
  explicit tvec4(T s) : x(s), y(s), z(s), w(s) {}
  tvec4(T a, T b, T c, T d) : x(a), y(b), z(c), w(d) {}
  tvec4& operator+=(const tvec4& o) { x += o.x; y += o.y; z += o.z; w += o.w; return *this; }
  bool operator==(const tvec4& o) const { return x == o.x && y == o.y && z == o.z && w == o.w; }
};

template <typename T> tvec3<T> operator+(tvec3<T> a, const tvec3<T>& b) { return a += b; }
template <typename T> tvec3<T> operator-(tvec3<T> a, const tvec3<T>& b) { return a -= b; }
template <typename T> tvec3<T> operator-(const tvec3<T>& a) { return {-a.x, -a.y, -a.z}; }
template <typename T> tvec3<T> operator*(const tvec3<T>& a, const tvec3<T>& b) {
  return {a.x * b.x, a.y * b.y, a.z * b.z};
}
template <typename T> tvec3<T> operator*(const tvec3<T>& a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> tvec3<T> operator*(T s, const tvec3<T>& a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> tvec3<T> operator/(const tvec3<T>& a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <typename T> tvec3<T> operator/(const tvec3<T>& a, const tvec3<T>& b) {
  return {a.x / b.x, a.y / b.y, a.z / b.z};
}
template <typename T> tvec3<T> operator+(const tvec3<T>& a, T s) { return {a.x + s, a.y + s, a.z + s}; }
template <typename T> tvec3<T> operator-(const tvec3<T>& a, T s) { return {a.x - s, a.y - s, a.z - s}; }
template <typename T> tvec4<T> operator+(tvec4<T> a, const tvec4<T>& b) { return a += b; }
template <typename T> tvec4<T> operator/(const tvec4<T>& a, T s) {
  return {a.x / s, a.y / s, a.z / s, a.w / s};
}

template <typename T> T dot(const tvec3<T>& a, const tvec3<T>& b) {
  return a.x * b.x + a.y * b.y + a.z * b.z;
}
template <typename T> T length(const tvec3<T>& a) { return std::sqrt(dot(a, a)); }
template <typename T> T distance(const tvec3<T>& a, const tvec3<T>& b) { return length(a - b); }
template <typename T> tvec3<T> normalize(const tvec3<T>& a) { return a * (T(1) / length(a)); }
template <typename T> tvec3<T> cross(const tvec3<T>& a, const tvec3<T>& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> tvec3<T> clamp(const tvec3<T>& a, T lo, T hi) {
  return {std::min(std::max(a.x, lo), hi), std::min(std::max(a.y, lo), hi),
          std::min(std::max(a.z, lo), hi)};
}
template <typename T> tvec3<T> min(const tvec3<T>& a, const tvec3<T>& b) {
  return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)};
}
template <typename T> tvec3<T> max(const tvec3<T>& a, const tvec3<T>& b) {
  return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)};
}
template <typename T> tvec3<T> floor(const tvec3<T>& a) {
  return {std::floor(a.x), std::floor(a.y), std::floor(a.z)};
}
/* triangle normal as glm/gtx/normal.hpp: normalize(cross(c - a, b - a)) */
template <typename T> tvec3<T> triangleNormal(const tvec3<T>& a, const tvec3<T>& b, const tvec3<T>& c) {
  return normalize(cross(c - a, b - a));
}
template <typename T> std::string to_string(const tvec3<T>& v) {
  return "vec3(" + std::to_string(v.x) + ", " + std::to_string(v.y) + ", " + std::to_string(v.z) + ")";
}

}  // namespace glm

#endif  // LIBFLUID_AMD_HAS_GLM
#endif  // LIBFLUID_AMD_VEC_HPP
