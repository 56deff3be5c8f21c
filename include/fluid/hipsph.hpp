/*
 * hip::SphSolver: the MI355X (gfx950) drop-in for libfluid's ocl::SphSolver
 * (include/fluid/oclsph.hpp:150-939) behind the same fluid::SphSolver<size_t,
 * float> interface (include/fluid/fluid.hpp:222-232).
 *
 *   // before                                   // after
 *   #include "fluid/oclsph.hpp"                 #include "fluid/hipsph.hpp"
 *   ocl::SphSolver s(0.1, kernelPath, dev);     hip::SphSolver s(0.1);
 *   auto r = s.advance(config, xs, colliders);  auto r = s.advance(config, xs, colliders);
 *
 * advance() keeps the reference's contract: sources add particles and drains
 * erase fluid particles on the host (oclsph.hpp:678-706); an empty set returns
 * an empty Result (:711-715); one step (predict, grid, colour diffusion,
 * iterations x lambda/delta, finalise) and the marching cubes of the
 * reference field run on the device; xs comes back permuted into this step's
 * Morton order (:874-892); queries report the fluid particles of the query
 * point's own cell (:770-791). Link with libfluid_amd/liblfamd.so.
 *
 * Mesh colliders (oclsph_collision.h) run inside every delta pass: the
 * triangles of all colliders, concatenated in order (oclsph.hpp:796-807).
 * Errors throw std::runtime_error carrying lf_last_error().
 */
#ifndef LIBFLUID_AMD_HIPSPH_HPP
#define LIBFLUID_AMD_HIPSPH_HPP

#include <cmath>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../libfluid_amd.h"
#include "fluid.hpp"

namespace hip {

// The records cross the C ABI without copies: same layouts.
using ParticleF = fluid::Particle<size_t, float>;
static_assert(sizeof(ParticleF) == sizeof(lf_particle),
              "fluid::Particle<size_t,float> must match lf_particle");
static_assert(offsetof(ParticleF, position) == offsetof(lf_particle, position),
              "particle position offset");
static_assert(offsetof(ParticleF, colour) == offsetof(lf_particle, colour),
              "particle colour offset");
static_assert(sizeof(geometry::MeshTriangle<float>) == sizeof(lf_mesh_triangle),
              "geometry::MeshTriangle<float> must match lf_mesh_triangle");
static_assert(sizeof(fluid::Type) == sizeof(int32_t), "fluid::Type is a 32-bit enum");

struct SphSolverOptions {
  float xsphC = 0.0f;     // XSPH viscosity (0: the reference has none)
  uint32_t flags = 0;     // extra LF_FLAG_* (LF_FLAG_KAHAN, ...)
  bool mesh = true;       // run the marching cubes (Result::triangles)
};

class SphSolver final : public fluid::SphSolver<size_t, float> {
 public:
  using Options = SphSolverOptions;

  explicit SphSolver(float h, int device = 0, Options options = Options()) : h(h), opt(options) {
    const int rc = lf_ctx_create(device, &ctx);
    if (rc != LF_OK)
      throw std::runtime_error("hip::SphSolver: no usable HIP device " + std::to_string(device) +
                               " (status " + std::to_string(rc) + ")");
  }
  ~SphSolver() override { lf_ctx_destroy(ctx); }
  SphSolver(const SphSolver&) = delete;
  SphSolver& operator=(const SphSolver&) = delete;

  fluid::Result<float> advance(const fluid::Config<float>& config,
                               std::vector<fluid::Particle<size_t, float>>& xs,
                               const std::vector<fluid::MeshCollider<float>>& colliders) override {
    std::vector<lf_triangle> mesh;  // oclsph.hpp:796-807
    for (const auto& c : colliders)
      for (const auto& t : c.triangles) {
        lf_triangle q;
        for (int k = 0; k < 3; ++k) {
          q.v0[k] = t.v0[k];
          q.v1[k] = t.v1[k];
          q.v2[k] = t.v2[k];
        }
        mesh.push_back(q);
      }
    check(lf_set_colliders(ctx, mesh.size(), mesh.data()));

    // sources then drains (oclsph.hpp:678-706)
    const float spacing = h * config.scale / 2;
    for (const fluid::Source<float>& source : config.sources) {
      const float size = std::sqrt(float(source.rate));
      const int width = int(std::floor(size)), depth = int(std::ceil(size));
      const auto offset = source.centre - (glm::tvec3<float>(float(width), 0.0f, float(depth)) /
                                           2.0f * spacing);
      for (int x = 0; x < width; ++x)
        for (int z = 0; z < depth; ++z)
          xs.emplace_back(source.tag, fluid::Type::Fluid, 1.0f, source.colour,
                          offset + glm::tvec3<float>(float(x), 0.0f, float(z)) * spacing,
                          source.velocity);
    }
    if (!config.drains.empty()) {
      std::vector<fluid::Particle<size_t, float>> kept;
      kept.reserve(xs.size());
      for (const auto& p : xs) {
        bool drop = false;
        if (p.type != fluid::Type::Obstacle)
          for (const auto& d : config.drains)
            if (glm::distance(d.centre, p.position) < d.width) drop = true;
        if (!drop) kept.push_back(p);
      }
      xs.swap(kept);
    }
    if (xs.empty()) return fluid::Result<float>({}, {});

    lf_pbf_params p;
    lf_pbf_params_default(&p, h);
    p.dt = config.dt;
    p.scale = config.scale;
    p.iterations = uint32_t(config.iteration);
    for (int k = 0; k < 3; ++k) {
      p.gravity[k] = config.constantForce[k];
      p.min_bound[k] = config.minBound[k];
      p.max_bound[k] = config.maxBound[k];
    }
    p.xsph_c = opt.xsphC;
    p.flags = opt.flags | LF_FLAG_DIFFUSE;

    std::vector<lf_well> wells(config.wells.size());
    for (size_t i = 0; i < wells.size(); ++i) {
      for (int k = 0; k < 3; ++k) wells[i].centre[k] = config.wells[i].centre[k];
      wells[i].force = config.wells[i].force;
    }
    check(lf_set_wells(ctx, wells.size(), wells.data()));

    lf_mc_params m{};  // ClMcConfig as advance() sets it (oclsph.hpp:669-673)
    m.field = LF_MC_FIELD_REFERENCE;
    m.isolevel = config.isolevel;
    m.resolution = config.resolution;
    m.particle_size = 25.0f;
    m.particle_influence = 0.5f;
    m.attributes = 1;

    std::vector<uint32_t> colourBefore(xs.size());
    for (size_t i = 0; i < xs.size(); ++i) colourBefore[i] = xs[i].colour;
    std::vector<uint32_t> order(xs.size());
    size_t ntri = 0;
    check(lf_advance(ctx, &p, opt.mesh ? &m : nullptr, reinterpret_cast<lf_particle*>(xs.data()),
                     xs.size(), nullptr, 0, &ntri, order.data()));
    fluid::Result<float> result;
    if (opt.mesh && ntri > 0) {
      result.triangles.resize(ntri);
      size_t got = 0;
      check(lf_mc_download(ctx, reinterpret_cast<lf_mesh_triangle*>(result.triangles.data()), ntri,
                           &got));
    }
    if (!config.queries.empty()) runQueries(config, xs, colourBefore, order, result);
    return result;
  }

  lf_ctx* context() const { return ctx; }

 private:
  static uint32_t spread10(uint32_t v) {
    v &= 1023u;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
  }
  static uint32_t cellOf(float q) {
    if (!(q >= 0.0f)) return 0u;
    if (q >= 4294967296.0f) return 1023u;
    return uint32_t(q) & 1023u;
  }

  // oclsph.hpp:770-791: fluid particles of the query point's own cell (the
  // colours they entered the step with, in this step's Morton order).
  void runQueries(const fluid::Config<float>& config,
                  const std::vector<fluid::Particle<size_t, float>>& xs,
                  const std::vector<uint32_t>& colourBefore, const std::vector<uint32_t>& order,
                  fluid::Result<float>& result) {
    uint32_t G = 0, extent[3];
    float minExt[3];
    check(lf_grid_info(ctx, &G, extent, minExt));
    std::vector<uint32_t> table(G);
    check(lf_read(ctx, LF_ARR_TABLE, table.data(), table.size() * sizeof(uint32_t)));
    for (const fluid::Query<float>& q : config.queries) {
      uint32_t c[3];
      for (int k = 0; k < 3; ++k) c[k] = cellOf((q.point[k] / config.scale - minExt[k]) / h);
      const uint32_t z = spread10(c[0]) | (spread10(c[1]) << 1) | (spread10(c[2]) << 2);
      if (!(size_t(z) + 1 < G)) continue;
      size_t N = 0;
      float avg[4] = {0, 0, 0, 0};
      for (uint32_t a = table[z]; a < table[z + 1]; ++a) {
        if (xs[a].type != fluid::Type::Fluid) continue;
        const uint32_t col = colourBefore[order[a]];
        ++N;
        for (int k = 0; k < 4; ++k) avg[k] += float((col >> (24 - 8 * k)) & 0xFFu);
      }
      if (N == 0) continue;
      uint32_t packed = 0;
      for (int k = 0; k < 4; ++k)
        packed = (packed << 8) | (uint32_t(uint8_t(avg[k] / float(N))) & 0xFFu);
      result.queries.emplace_back(q.id, q.point, N, packed);
    }
  }

  void check(int rc) const {
    if (rc != LF_OK)
      throw std::runtime_error(std::string("hip::SphSolver: ") + lf_last_error(ctx) +
                               " (status " + std::to_string(rc) + ")");
  }

  float h;
  Options opt;
  lf_ctx* ctx = nullptr;
};

}  // namespace hip

#endif  // LIBFLUID_AMD_HIPSPH_HPP
