// Exercises include/fluid/wire.hpp from C++ for tests/test_wire.py:
//   wire_main write DIR   writes scene.mmf, particles.mmf, triangles.mmf,
//                         query.mmf, colliders.mmf, dynamic_bodies.mmf and
//                         defs.json into DIR (fixed contents, below)
//   wire_main read FILE   reads FILE as scene.mmf and prints its fields
// and checks, before either, that every field width in wire::layouts() is
// sizeof the member it stands for (the reference's mmf::Entry::size).
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "fluid/fluid.hpp"
#include "fluid/geometry.hpp"
#include "fluid/wire.hpp"

using P = fluid::Particle<size_t, float>;
using MT = geometry::MeshTriangle<float>;

#define SZ(T, m) sizeof(decltype(std::declval<T>().m))

static bool checkWidths() {
  using namespace fluid;
  const std::vector<std::pair<std::string, std::vector<size_t>>> members = {
      {"header", {sizeof(int64_t), sizeof(size_t), sizeof(size_t)}},  // long, size_t, size_t
      {"particle",
       {SZ(P, id), SZ(P, type), SZ(P, mass), SZ(P, colour), SZ(P, position.x), SZ(P, position.y),
        SZ(P, position.z), SZ(P, velocity.x), SZ(P, velocity.y), SZ(P, velocity.z)}},
      {"well", {SZ(Well<float>, centre.x), SZ(Well<float>, centre.y), SZ(Well<float>, centre.z),
                SZ(Well<float>, force)}},
      {"source",
       {SZ(Source<float>, centre.x), SZ(Source<float>, centre.y), SZ(Source<float>, centre.z),
        SZ(Source<float>, velocity.x), SZ(Source<float>, velocity.y),
        SZ(Source<float>, velocity.z), SZ(Source<float>, rate), SZ(Source<float>, tag),
        SZ(Source<float>, colour)}},
      {"drain", {SZ(Drain<float>, centre.x), SZ(Drain<float>, centre.y),
                 SZ(Drain<float>, centre.z), SZ(Drain<float>, width), SZ(Drain<float>, depth)}},
      {"query", {SZ(Query<float>, id), SZ(Query<float>, point.x), SZ(Query<float>, point.y),
                 SZ(Query<float>, point.z)}},
      {"queryResult",
       {SZ(QueryResult<float>, id), SZ(QueryResult<float>, point.x),
        SZ(QueryResult<float>, point.y), SZ(QueryResult<float>, point.z),
        SZ(QueryResult<float>, neighbours), SZ(QueryResult<float>, averageColour)}},
      {"sceneMeta",
       {sizeof(bool), sizeof(bool), sizeof(int), 4, 4, 4, 4, 4, 4, 4, 4, 4, 4}},
      {"vec3", {sizeof(float), sizeof(float), sizeof(float)}},
      {"triangle", std::vector<size_t>(9, SZ(geometry::Triangle<float>, v0.x))},
      {"meshTriangle",
       {SZ(MT, v0.x), SZ(MT, v0.y), SZ(MT, v0.z), SZ(MT, v1.x), SZ(MT, v1.y), SZ(MT, v1.z),
        SZ(MT, v2.x), SZ(MT, v2.y), SZ(MT, v2.z), SZ(MT, n0.x), SZ(MT, n0.y), SZ(MT, n0.z),
        SZ(MT, n1.x), SZ(MT, n1.y), SZ(MT, n1.z), SZ(MT, n2.x), SZ(MT, n2.y), SZ(MT, n2.z),
        SZ(MT, c.x), SZ(MT, c.y), SZ(MT, c.z)}},
  };
  bool ok = members.size() == wire::layouts().size();
  for (const auto& m : members) {
    const wire::Layout& l = wire::layout(m.first);
    if (l.fields.size() != m.second.size()) {
      std::fprintf(stderr, "%s: %zu fields, %zu members\n", m.first.c_str(), l.fields.size(),
                   m.second.size());
      ok = false;
      continue;
    }
    for (size_t i = 0; i < m.second.size(); ++i)
      if (l.fields[i].size != m.second[i]) {
        std::fprintf(stderr, "%s.%s: width %zu, member %zu\n", m.first.c_str(),
                     l.fields[i].name, l.fields[i].size, m.second[i]);
        ok = false;
      }
  }
  return ok;
}

static int write(const std::string& dir) {
  wire::Scene<float> s;
  s.meta.suspend = false;
  s.meta.terminate = true;
  s.meta.solverIter = 5;
  s.meta.solverStep = 1.5f;
  s.meta.solverScale = 1000.0f;
  s.meta.surfaceRes = 1.0f;
  s.meta.gravity = 9.8f;
  s.meta.minBound = glm::tvec3<float>(0.0f, -1.0f, 2.0f);
  s.meta.maxBound = glm::tvec3<float>(1000.0f, 900.0f, 800.0f);
  s.wells = {{glm::tvec3<float>(1, 2, 3), 4.5f}, {glm::tvec3<float>(-1, -2, -3), -4.5f}};
  fluid::Source<float> src;
  src.centre = glm::tvec3<float>(10, 20, 30);
  src.velocity = glm::tvec3<float>(0.5f, -0.25f, 0.125f);
  src.rate = 18;
  src.tag = 0x123456789ull;
  src.colour = 0xFF00FF00u;
  s.sources = {src};
  s.drains = {{glm::tvec3<float>(7, 8, 9), 0.75f, 0.5f}};
  s.queries = {{11, glm::tvec3<float>(1, 1, 1)}, {12, glm::tvec3<float>(2, 2, 2)},
               {13, glm::tvec3<float>(3, 3, 3)}};
  {
    auto f = wire::MappedFile::createSink(dir + "/scene.mmf", 1 << 16);
    wire::writeScene(f.sink(), s);
  }
  std::vector<P> ps;
  for (size_t i = 0; i < 5; ++i)
    ps.emplace_back(1000 + i, i % 2 ? fluid::Obstacle : fluid::Fluid, 1.0f + float(i),
                    0xA0B0C0D0u + uint32_t(i), glm::tvec3<float>(float(i), 2.0f * i, -1.0f * i),
                    glm::tvec3<float>(0.5f * i, 0.25f, -0.125f));
  {
    auto f = wire::MappedFile::createSink(dir + "/particles.mmf",
                                          wire::kHeaderBytes + 10 * wire::kParticleBytes);
    wire::writeParticles(f.sink(), ps);
  }
  std::vector<MT> ts;
  for (int i = 0; i < 3; ++i) {
    MT t(glm::tvec3<float>(i, 0, 0), glm::tvec3<float>(0, i, 0), glm::tvec3<float>(0, 0, i),
         glm::tvec3<float>(1, 0, 0), glm::tvec3<float>(0, 1, 0), glm::tvec3<float>(0, 0, 1));
    t.c = glm::tvec3<uint32_t>(0x11111111u * uint32_t(i + 1), 2u, 3u);
    ts.push_back(t);
  }
  {
    auto f = wire::MappedFile::createSink(dir + "/triangles.mmf", 1 << 14);
    wire::writeTriangles(f.sink(), ts);
  }
  std::vector<fluid::QueryResult<float>> qs = {
      fluid::QueryResult<float>(11, glm::tvec3<float>(1, 1, 1), 27, 0x80402010u),
      fluid::QueryResult<float>(13, glm::tvec3<float>(3, 3, 3), 5, 0xFFFFFFFFu)};
  {
    auto f = wire::MappedFile::createSink(dir + "/query.mmf", 1 << 12);
    wire::writeQueries(f.sink(), qs);
  }
  {
    std::vector<geometry::Triangle<float>> tri = {
        geometry::Triangle<float>(glm::tvec3<float>(0, 0, 0), glm::tvec3<float>(1, 0, 0),
                                  glm::tvec3<float>(0, 1, 0))};
    auto f = wire::MappedFile::createSink(dir + "/colliders.mmf", 1 << 12);
    wire::writeBlock(f.sink(), tri);
    std::vector<glm::tvec3<float>> body = {glm::tvec3<float>(1, 2, 3), glm::tvec3<float>(4, 5, 6)};
    auto g = wire::MappedFile::createSink(dir + "/dynamic_bodies.mmf", 1 << 12);
    wire::Writer w(g.sink(), 0);
    wire::Header h(body.size());
    h.written = body.size();
    wire::put(w, h);
    for (const auto& v : body) w.put3(v);
  }
  std::ofstream(dir + "/defs.json") << wire::defsJson();
  // read back through the C++ readers
  auto sc = wire::MappedFile::createSource(dir + "/scene.mmf");
  const wire::Scene<float> back = wire::readScene<float>(sc.source());
  auto pf = wire::MappedFile::createSource(dir + "/particles.mmf");
  const auto pback = wire::readParticles<size_t, float>(pf.source());
  auto cf = wire::MappedFile::createSource(dir + "/colliders.mmf");
  const auto col = wire::readCollider<float>(cf.source());
  auto bf = wire::MappedFile::createSource(dir + "/dynamic_bodies.mmf");
  const auto rb = wire::readRigidBody<float>(bf.source());
  const bool same = back.wells.size() == 2 && back.sources.size() == 1 &&
                    back.sources[0].tag == src.tag && back.drains.size() == 1 &&
                    back.queries.size() == 3 && back.queries[2].id == 13 && pback == ps &&
                    col.triangles.size() == 1 && col.triangles[0].v1.x == 1.0f &&
                    rb.obstacles.size() == 2 && rb.obstacles[1].z == 6.0f;
  std::printf("roundtrip %s\n", same ? "ok" : "MISMATCH");
  return same ? 0 : 1;
}

static int read(const std::string& path) {
  auto f = wire::MappedFile::createSource(path);
  const wire::Scene<float> s = wire::readScene<float>(f.source());
  std::printf("meta %d %d %d %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g %.9g\n", s.meta.suspend,
              s.meta.terminate, s.meta.solverIter, s.meta.solverStep, s.meta.solverScale,
              s.meta.surfaceRes, s.meta.gravity, s.meta.minBound.x, s.meta.minBound.y,
              s.meta.minBound.z, s.meta.maxBound.x, s.meta.maxBound.y, s.meta.maxBound.z);
  for (const auto& w : s.wells)
    std::printf("well %.9g %.9g %.9g %.9g\n", w.centre.x, w.centre.y, w.centre.z, w.force);
  for (const auto& x : s.sources)
    std::printf("source %.9g %.9g %.9g %.9g %.9g %.9g %zu %zu %u\n", x.centre.x, x.centre.y,
                x.centre.z, x.velocity.x, x.velocity.y, x.velocity.z, x.rate, x.tag, x.colour);
  for (const auto& d : s.drains)
    std::printf("drain %.9g %.9g %.9g %.9g %.9g\n", d.centre.x, d.centre.y, d.centre.z, d.width,
                d.depth);
  for (const auto& q : s.queries)
    std::printf("query %zu %.9g %.9g %.9g\n", q.id, q.point.x, q.point.y, q.point.z);
  return 0;
}

int main(int argc, char** argv) {
  if (!checkWidths()) {
    std::fprintf(stderr, "field widths differ from the member sizes\n");
    return 2;
  }
  if (argc == 3 && std::string(argv[1]) == "write") return write(argv[2]);
  if (argc == 3 && std::string(argv[1]) == "read") return read(argv[2]);
  std::fprintf(stderr, "usage: wire_main write DIR | read FILE\n");
  return 2;
}
