/*
 * The fluid:: solver API of libfluid (include/fluid/fluid.hpp:13-234),
 * source-compatible: the same records, members, constructors and the abstract
 * SphSolver<T,N>::advance, so code written against libfluid compiles against
 * this header and can swap ocl::SphSolver for hip::SphSolver
 * (include/fluid/hipsph.hpp).
 */
#ifndef LIBFLUID_AMD_FLUID_HPP
#define LIBFLUID_AMD_FLUID_HPP

#include <cstdint>
#include <ostream>
#include <vector>

#include "geometry.hpp"
#include "vec.hpp"

namespace fluid {

using glm::tvec3;
using glm::tvec4;

enum Type { Fluid = 0, Obstacle = 1 };

template <typename N> struct Query {
  size_t id;
  tvec3<N> point;
};

template <typename N> struct QueryResult {
  size_t id;
  tvec3<N> point;
  size_t neighbours;
  uint32_t averageColour;
  explicit QueryResult(size_t id, tvec3<N> point, size_t neighbours, uint32_t averageColour)
      : id(id), point(point), neighbours(neighbours), averageColour(averageColour) {}
};

template <typename T, typename N> struct Particle {
  T id;
  Type type;
  N mass;
  tvec3<N> position;
  tvec3<N> velocity;
  uint32_t colour;
  Particle() : type(Type::Fluid) {}
  explicit Particle(T t, Type type, N mass, const uint32_t& colour, const tvec3<N>& position,
                    const tvec3<N>& velocity)
      : id(t), type(type), mass(mass), position(position), velocity(velocity), colour(colour) {}
  bool operator==(const Particle& r) const {
    return id == r.id && type == r.type && mass == r.mass && colour == r.colour &&
           position == r.position && velocity == r.velocity;
  }
  bool operator!=(const Particle& r) const { return !(r == *this); }
  friend std::ostream& operator<<(std::ostream& os, const Particle& p) {
    return os << "t: " << p.id << ", mass: " << p.mass << ", colour: 0x" << std::hex << p.colour
              << std::dec << ", pos: " << glm::to_string(p.position)
              << ", vel: " << glm::to_string(p.velocity);
  }
};

template <typename N> struct Ray {
  Ray(const tvec3<N>& prev, const tvec3<N>& origin, const tvec3<N>& velocity)
      : prev(prev), origin(origin), velocity(velocity) {}
  const tvec3<N>& getPrev() const { return prev; }
  const tvec3<N>& getOrigin() const { return origin; }
  const tvec3<N>& getVelocity() const { return velocity; }

 private:
  tvec3<N> prev, origin, velocity;
};

template <typename N> struct Response {
  Response(const tvec3<N>& position, const tvec3<N>& velocity)
      : position(position), velocity(velocity) {}
  const tvec3<N>& getPosition() const { return position; }
  const tvec3<N>& getVelocity() const { return velocity; }

 private:
  tvec3<N> position, velocity;
};

template <typename N> struct MeshCollider {
  const std::vector<geometry::Triangle<N>> triangles;
  explicit MeshCollider(const std::vector<geometry::Triangle<N>>& triangles)
      : triangles(triangles) {}
};

template <typename N> struct RigidBody {
  const std::vector<tvec3<N>> obstacles;
  explicit RigidBody(const std::vector<tvec3<N>>& obstacles) : obstacles(obstacles) {}
};

template <typename N> struct Well {
  tvec3<N> centre;
  N force;
};

template <typename N> struct Source {
  tvec3<N> centre;
  tvec3<N> velocity;
  size_t rate, tag;
  uint32_t colour;
};

template <typename N> struct Drain {
  tvec3<N> centre;
  N width, depth;
};

template <typename N> struct Config {
  const N dt;
  const N scale;
  const N resolution;
  const N isolevel;
  const size_t iteration;
  const tvec3<N> constantForce;
  const std::vector<fluid::Well<N>> wells;
  const std::vector<fluid::Source<N>> sources;
  const std::vector<fluid::Drain<N>> drains;
  const std::vector<fluid::Query<N>> queries;
  const tvec3<N> minBound, maxBound;
  explicit Config(N dt, N scale, N resolution, N isolevel, size_t iteration,
                  const tvec3<N>& constantForce, const std::vector<fluid::Well<N>>& wells,
                  const std::vector<fluid::Source<N>>& sources,
                  const std::vector<fluid::Drain<N>>& drains,
                  const std::vector<fluid::Query<N>>& queries, const tvec3<N>& min,
                  const tvec3<N>& max)
      : dt(dt), scale(scale), resolution(resolution), isolevel(isolevel), iteration(iteration),
        constantForce(constantForce), wells(wells), sources(sources), drains(drains),
        queries(queries), minBound(min), maxBound(max) {}
};

template <typename N> struct Result {
  std::vector<geometry::MeshTriangle<N>> triangles{};
  std::vector<QueryResult<N>> queries{};
  Result() = default;
  Result(const std::vector<geometry::MeshTriangle<N>>& triangles,
         const std::vector<QueryResult<N>>& queries)
      : triangles(triangles), queries(queries) {}
};

template <typename T, typename N> class SphSolver {
 public:
  virtual ~SphSolver() = default;
  virtual Result<N> advance(const Config<N>& config, std::vector<Particle<T, N>>& xs,
                            const std::vector<MeshCollider<N>>& colliders) = 0;
};

}  // namespace fluid

#endif  // LIBFLUID_AMD_FLUID_HPP
