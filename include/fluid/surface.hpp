/*
 * surface:: CPU marching cubes of libfluid (include/fluid/surface.hpp:20-387),
 * source-compatible: Lattice<T>, GridCell<N>, lerp, polygonise,
 * marchSingleN, marchSingle, createLattice and parameterise, on the tables of
 * mctable.hpp. The GPU path is lf_mc_* (libfluid_amd.h); this header is the
 * host-side mesher the reference also ships.
 *
 * Semantics kept from the reference: corner bit i set when value_i <
 * isolevel (CUBE_OFFSETS order); lerp snaps to an endpoint when |iso - v| or
 * |v1 - v0| is below 1e-5 (surface.hpp:48-53). Triangles are appended in
 * cube order (x, then y, then z) and, within a cube, in table order; the
 * reference's OpenMP vector merge leaves that order unspecified.
 */
#ifndef LIBFLUID_AMD_SURFACE_HPP
#define LIBFLUID_AMD_SURFACE_HPP

#include <array>
#include <cmath>
#include <cstddef>
#include <functional>
#include <vector>

#include "geometry.hpp"
#include "mctable.hpp"

namespace surface {

using geometry::MeshTriangle;
using glm::tvec3;

const static std::array<tvec3<size_t>, 8> CUBE_OFFSETS = {
    tvec3<size_t>(0, 0, 0), tvec3<size_t>(1, 0, 0), tvec3<size_t>(1, 1, 0), tvec3<size_t>(0, 1, 0),
    tvec3<size_t>(0, 0, 1), tvec3<size_t>(1, 0, 1), tvec3<size_t>(1, 1, 1), tvec3<size_t>(0, 1, 1)};

template <typename N> struct GridCell {
  std::array<tvec3<N>, 8> p;
  std::array<N, 8> val;
  explicit GridCell(const std::array<tvec3<N>, 8>& p, const std::array<N, 8>& val)
      : p(p), val(val) {}
};

template <typename N>
inline tvec3<N> lerp(N isolevel, tvec3<N> p1, tvec3<N> p2, N v1, N v2) {
  if (std::abs(isolevel - v1) < N(0.00001)) return p1;
  if (std::abs(isolevel - v2) < N(0.00001)) return p2;
  if (std::abs(v1 - v2) < N(0.00001)) return p1;
  return p1 + (p2 - p1) * ((isolevel - v1) / (v2 - v1));
}

namespace detail {
template <typename N> inline size_t caseIndex(N isolevel, const std::array<N, 8>& v) {
  size_t ci = 0;
  for (int i = 0; i < 8; ++i)
    if (v[i] < isolevel) ci |= size_t(1) << i;
  return ci;
}
// Interpolated vertex of every cut edge of the case.
template <typename N, typename P>
inline void cutEdges(N isolevel, size_t ci, const std::array<N, 8>& v, const std::array<P, 8>& p,
                     std::array<P, 12>& out) {
  const int edges = mctable::EdgeTable[ci];
  for (int e = 0; e < 12; ++e)
    if (edges & (1 << e)) {
      const int a = mctable::kEdgeCorners[e][0], b = mctable::kEdgeCorners[e][1];
      out[e] = lerp(isolevel, p[a], p[b], v[a], v[b]);
    }
}
}  // namespace detail

// http://paulbourke.net/geometry/polygonise/ (surface.hpp:62-98)
template <typename N>
void polygonise(const GridCell<N>& g, const N isolevel, std::vector<MeshTriangle<N>>& sink) {
  const size_t ci = detail::caseIndex(isolevel, g.val);
  if (mctable::EdgeTable[ci] == 0) return;
  std::array<tvec3<N>, 12> vs;
  detail::cutEdges(isolevel, ci, g.val, g.p, vs);
  for (size_t i = 0; mctable::TriTable[ci][i] != -1; i += 3)
    sink.emplace_back(vs[mctable::TriTable[ci][i]], vs[mctable::TriTable[ci][i + 1]],
                      vs[mctable::TriTable[ci][i + 2]]);
}

template <typename T> class Lattice {
  const size_t d1, d2, d3;
  std::vector<T> data;

 public:
  explicit Lattice(const size_t d1 = 0, const size_t d2 = 0, const size_t d3 = 0,
                   T const& t = T())
      : d1(d1), d2(d2), d3(d3), data(d1 * d2 * d3, t) {}
  T& operator()(size_t i, size_t j, size_t k) { return data[i * d2 * d3 + j * d3 + k]; }
  T const& operator()(size_t i, size_t j, size_t k) const { return data[i * d2 * d3 + j * d3 + k]; }
  inline size_t xSize() const { return d1; }
  inline size_t ySize() const { return d2; }
  inline size_t zSize() const { return d3; }
  inline size_t size() const { return data.size(); }
  decltype(auto) begin() { return data.begin(); }
  decltype(auto) end() { return data.end(); }
};

template <typename N> using MCLattice = Lattice<tvec3<N>>;

template <typename N>
MCLattice<N> createLattice(const size_t xSize, const size_t ySize, const size_t zSize,
                           const N offset, const N factor) {
  MCLattice<N> lattice(xSize, ySize, zSize, tvec3<N>(0));
  for (size_t x = 0; x < xSize; ++x)
    for (size_t y = 0; y < ySize; ++y)
      for (size_t z = 0; z < zSize; ++z)
        lattice(x, y, z) = tvec3<N>(static_cast<N>(x) * factor + offset,
                                    static_cast<N>(y) * factor + offset,
                                    static_cast<N>(z) * factor + offset);
  return lattice;
}

// Vertex count of a cube's case, 0 when the cube is not cut (surface.hpp:166-182).
template <typename N>
inline unsigned char marchSingleN(const N isolevel, const std::array<N, 8>& vertices) {
  const size_t ci = detail::caseIndex(isolevel, vertices);
  if (mctable::EdgeTable[ci] == 0) return 0;
  return mctable::NumVertsTable[ci];
}

// One cube with per-corner normals (surface.hpp:185-266).
template <typename N>
inline void marchSingle(const N isolevel, const std::array<N, 8>& vertices,
                        const std::array<tvec3<N>, 8>& normals,
                        const std::array<tvec3<N>, 8>& pos,
                        std::vector<MeshTriangle<N>>& triangles) {
  const size_t ci = detail::caseIndex(isolevel, vertices);
  if (mctable::EdgeTable[ci] == 0) return;
  std::array<tvec3<N>, 12> ts, ns;
  detail::cutEdges(isolevel, ci, vertices, pos, ts);
  detail::cutEdges(isolevel, ci, vertices, normals, ns);
  for (size_t i = 0; mctable::TriTable[ci][i] != -1; i += 3) {
    const int x = mctable::TriTable[ci][i], y = mctable::TriTable[ci][i + 1],
              z = mctable::TriTable[ci][i + 2];
    triangles.emplace_back(ts[x], ts[y], ts[z], ns[x], ns[y], ns[z]);
  }
}

// Marching cubes of f sampled at the lattice points (surface.hpp:269-384).
template <typename N>
std::vector<MeshTriangle<N>> parameterise(const N isolevel, MCLattice<N> lattice,
                                          const std::function<N(tvec3<N>&)>& f) {
  Lattice<N> field(lattice.xSize(), lattice.ySize(), lattice.zSize(), 0);
#pragma omp parallel for
  for (long x = 0; x < long(lattice.xSize()); ++x)
    for (size_t y = 0; y < lattice.ySize(); ++y)
      for (size_t z = 0; z < lattice.zSize(); ++z) field(size_t(x), y, z) = f(lattice(size_t(x), y, z));
  std::vector<MeshTriangle<N>> triangles;
  if (lattice.xSize() < 2 || lattice.ySize() < 2 || lattice.zSize() < 2) return triangles;
  for (size_t x = 0; x + 1 < lattice.xSize(); ++x)
    for (size_t y = 0; y + 1 < lattice.ySize(); ++y)
      for (size_t z = 0; z + 1 < lattice.zSize(); ++z) {
        std::array<N, 8> ns{};
        std::array<tvec3<N>, 8> vs{};
        for (size_t j = 0; j < 8; ++j) {
          const tvec3<size_t> o = tvec3<size_t>(x, y, z) + CUBE_OFFSETS[j];
          vs[j] = lattice(o.x, o.y, o.z);
          ns[j] = field(o.x, o.y, o.z);
        }
        polygonise(GridCell<N>(vs, ns), isolevel, triangles);
      }
  return triangles;
}

}  // namespace surface

#endif  // LIBFLUID_AMD_SURFACE_HPP
