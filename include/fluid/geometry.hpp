/*
 * geometry:: records of libfluid's API (include/fluid/geometry.hpp:17-80):
 * Triangle<N> and MeshTriangle<N> with the same members and layout, so
 * Result<float>::triangles is an array of lf_mesh_triangle (84 bytes each for
 * N = float).
 */
#ifndef LIBFLUID_AMD_GEOMETRY_HPP
#define LIBFLUID_AMD_GEOMETRY_HPP

#include <cstdint>
#include <ostream>

#include "vec.hpp"

namespace geometry {

using glm::tvec3;

template <typename N> struct Triangle {
  tvec3<N> v0, v1, v2;
  Triangle() = default;
  explicit Triangle(const tvec3<N>& v0, const tvec3<N>& v1, const tvec3<N>& v2)
      : v0(v0), v1(v1), v2(v2) {}
  friend std::ostream& operator<<(std::ostream& os, const Triangle& t) {
    return os << "T(" << glm::to_string(t.v0) << "," << glm::to_string(t.v1) << ","
              << glm::to_string(t.v2) << ")";
  }
};

template <typename N> struct MeshTriangle {
  tvec3<N> v0, v1, v2;
  tvec3<N> n0, n1, n2;
  tvec3<uint32_t> c;
  MeshTriangle() = default;
  explicit MeshTriangle(const tvec3<N>& v0, const tvec3<N>& v1, const tvec3<N>& v2,
                        const tvec3<N>& normal)
      : v0(v0), v1(v1), v2(v2), n0(normal), n1(normal), n2(normal),
        c(tvec3<uint32_t>(0xFFFFFFFFu)) {}
  explicit MeshTriangle(const tvec3<N>& v0, const tvec3<N>& v1, const tvec3<N>& v2)
      : v0(v0), v1(v1), v2(v2), n0(glm::triangleNormal(v0, v1, v2)), n1(n0), n2(n0) {}
  explicit MeshTriangle(const tvec3<N>& v0, const tvec3<N>& v1, const tvec3<N>& v2,
                        const tvec3<N>& n0, const tvec3<N>& n1, const tvec3<N>& n2)
      : v0(v0), v1(v1), v2(v2), n0(n0), n1(n1), n2(n2) {}
  friend std::ostream& operator<<(std::ostream& os, const MeshTriangle& t) {
    return os << "MT(v=<" << glm::to_string(t.v0) << "," << glm::to_string(t.v1) << ","
              << glm::to_string(t.v2) << "> n=<" << glm::to_string(t.n0) << ", "
              << glm::to_string(t.n1) << ", " << glm::to_string(t.n2) << "> c=<0x" << std::hex
              << t.c.x << ", 0x" << t.c.y << ", 0x" << t.c.z << std::dec << ">)";
  }
};

}  // namespace geometry

namespace surface {
using geometry::MeshTriangle;  // libfluid also reaches it as surface::MeshTriangle
}

#endif  // LIBFLUID_AMD_GEOMETRY_HPP
