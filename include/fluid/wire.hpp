/*
 * The memory-mapped file wire format libfluid's sample application uses to
 * talk to its external renderer (include/fluid/mmf.hpp:10-187,
 * samples/structures.hpp:11-508), restated without the mmf/mio/nlohmann
 * template machinery:
 *
 *  - every record is its fields packed back to back, little-endian, each
 *    field as many bytes as its C++ type (no padding, no alignment);
 *  - a block is a 24-byte header {int64 timestamp ms, u64 entries, u64 written}
 *    followed by `entries` records; the writer stores the header with
 *    written = 0, then the records, then the header again with
 *    written = entries, so a reader treats written != entries as "in
 *    progress" (structures.hpp:391-425);
 *  - scene.mmf is SceneMeta followed by four blocks (wells, sources, drains,
 *    queries; structures.hpp:432-476);
 *  - defs.json lists every record's fields and sizes (samples/main.cpp:146-182).
 *
 * Field order and sizes per record (T = size_t, N = float):
 *   header        timestamp 8, entries 8, written 8                      = 24
 *   particle      id 8, type 4, mass 4, colour 4, position.xyz, velocity.xyz = 44
 *   meshTriangle  v0..v2 xyz, n0..n2 xyz, c.xyz (u32)                    = 84
 *   queryResult   query.id 8, query.xyz, neighbours 8, avgcolour 4       = 32
 *   query         query.id 8, query.xyz                                  = 20
 *   well          well.xyz, force                                        = 16
 *   source        source.xyz, source.vel.xyz, rate 8, tag 8, colour 4    = 44
 *   drain         drain.xyz, width, depth                                = 20
 *   triangle      v0..v2 xyz                                             = 36
 *   vec3          vec3.xyz                                               = 12
 *   sceneMeta     suspend 1, terminate 1, solverIter 4, solverStep,
 *                 solverScale, surfaceRes, gravity, minBound.xyz,
 *                 maxBound.xyz                                           = 46
 *
 * The device-side writers lf_wire_particles / lf_wire_triangles
 * (libfluid_amd.h) pack particle and triangle blocks on the GPU from the
 * context's state and copy them into the same layout.
 */
#ifndef LIBFLUID_AMD_WIRE_HPP
#define LIBFLUID_AMD_WIRE_HPP

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "fluid.hpp"
#include "geometry.hpp"

namespace wire {

// ---- record layouts ---------------------------------------------------------

struct Field {
  const char* name;
  size_t size;
};

struct Layout {
  const char* type;
  std::vector<Field> fields;
  size_t size() const {
    size_t s = 0;
    for (const Field& f : fields) s += f.size;
    return s;
  }
};

// Every record the reference's defs.json describes, keyed as there.
inline const std::vector<Layout>& layouts() {
  static const std::vector<Layout> all = {
      {"header", {{"timestamp", 8}, {"entries", 8}, {"written", 8}}},
      {"well", {{"well.x", 4}, {"well.y", 4}, {"well.z", 4}, {"force", 4}}},
      {"source",
       {{"source.x", 4}, {"source.y", 4}, {"source.z", 4}, {"source.vel.x", 4},
        {"source.vel.y", 4}, {"source.vel.z", 4}, {"rate", 8}, {"tag", 8}, {"colour", 4}}},
      {"drain", {{"drain.x", 4}, {"drain.y", 4}, {"drain.z", 4}, {"width", 4}, {"depth", 4}}},
      {"sceneMeta",
       {{"suspend", 1}, {"terminate", 1}, {"solverIter", 4}, {"solverStep", 4},
        {"solverScale", 4}, {"surfaceRes", 4}, {"gravity", 4}, {"minBound.x", 4},
        {"minBound.y", 4}, {"minBound.z", 4}, {"maxBound.x", 4}, {"maxBound.y", 4},
        {"maxBound.z", 4}}},
      {"vec3", {{"vec3.x", 4}, {"vec3.y", 4}, {"vec3.z", 4}}},
      {"query", {{"query.id", 8}, {"query.x", 4}, {"query.y", 4}, {"query.z", 4}}},
      {"queryResult",
       {{"query.id", 8}, {"query.x", 4}, {"query.y", 4}, {"query.z", 4}, {"neighbours", 8},
        {"avgcolour", 4}}},
      {"particle",
       {{"id", 8}, {"type", 4}, {"mass", 4}, {"colour", 4}, {"position.x", 4},
        {"position.y", 4}, {"position.z", 4}, {"velocity.x", 4}, {"velocity.y", 4},
        {"velocity.z", 4}}},
      {"triangle",
       {{"v0.x", 4}, {"v0.y", 4}, {"v0.z", 4}, {"v1.x", 4}, {"v1.y", 4}, {"v1.z", 4},
        {"v2.x", 4}, {"v2.y", 4}, {"v2.z", 4}}},
      {"meshTriangle",
       {{"v0.x", 4}, {"v0.y", 4}, {"v0.z", 4}, {"v1.x", 4}, {"v1.y", 4}, {"v1.z", 4},
        {"v2.x", 4}, {"v2.y", 4}, {"v2.z", 4}, {"n0.x", 4}, {"n0.y", 4}, {"n0.z", 4},
        {"n1.x", 4}, {"n1.y", 4}, {"n1.z", 4}, {"n2.x", 4}, {"n2.y", 4}, {"n2.z", 4},
        {"c.x", 4}, {"c.y", 4}, {"c.z", 4}}},
  };
  return all;
}

inline const Layout& layout(const std::string& type) {
  for (const Layout& l : layouts())
    if (type == l.type) return l;
  throw std::invalid_argument("wire: no record type " + type);
}

// defs.json as the sample writes it (nlohmann::json::dump(1): one-space
// indent, object keys in sorted order).
inline std::string defsJson() {
  std::vector<const Layout*> sorted;
  for (const Layout& l : layouts()) sorted.push_back(&l);
  for (size_t i = 1; i < sorted.size(); ++i)
    for (size_t j = i; j > 0 && std::strcmp(sorted[j - 1]->type, sorted[j]->type) > 0; --j)
      std::swap(sorted[j - 1], sorted[j]);
  std::string s = "{\n";
  for (size_t k = 0; k < sorted.size(); ++k) {
    const Layout& l = *sorted[k];
    s += " \"" + std::string(l.type) + "\": {\n  \"fields\": [\n";
    for (size_t f = 0; f < l.fields.size(); ++f) {
      s += "   {\n    \"name\": \"" + std::string(l.fields[f].name) + "\",\n    \"size\": " +
           std::to_string(l.fields[f].size) + "\n   }";
      s += f + 1 < l.fields.size() ? ",\n" : "\n";
    }
    s += "  ],\n  \"size\": " + std::to_string(l.size()) + "\n }";
    s += k + 1 < sorted.size() ? ",\n" : "\n";
  }
  return s + "}";
}

constexpr size_t kHeaderBytes = 24;
constexpr size_t kParticleBytes = 44;
constexpr size_t kMeshTriangleBytes = 84;
constexpr size_t kQueryResultBytes = 32;

// ---- byte spans ---------------------------------------------------------------

// A writable window onto a mapped file (or any buffer). Every access is
// bounds-checked: a record that does not fit throws instead of writing.
struct Sink {
  unsigned char* data = nullptr;
  size_t size = 0;
};

struct Source {
  const unsigned char* data = nullptr;
  size_t size = 0;
};

class Writer {
 public:
  Writer(Sink s, size_t offset) : s_(s), off_(offset) {}
  template <typename V> Writer& put(const V& v) {
    if (off_ + sizeof(V) > s_.size) throw std::out_of_range("wire: write past the end of the sink");
    std::memcpy(s_.data + off_, &v, sizeof(V));
    off_ += sizeof(V);
    return *this;
  }
  template <typename V> Writer& put3(const glm::tvec3<V>& v) { return put(v.x).put(v.y).put(v.z); }
  size_t offset() const { return off_; }

 private:
  Sink s_;
  size_t off_;
};

class Reader {
 public:
  Reader(Source s, size_t offset) : s_(s), off_(offset) {}
  template <typename V> V get() {
    if (off_ + sizeof(V) > s_.size) throw std::out_of_range("wire: read past the end of the source");
    V v;
    std::memcpy(&v, s_.data + off_, sizeof(V));
    off_ += sizeof(V);
    return v;
  }
  template <typename V> glm::tvec3<V> get3() {
    const V x = get<V>(), y = get<V>(), z = get<V>();
    return glm::tvec3<V>(x, y, z);
  }
  size_t offset() const { return off_; }

 private:
  Source s_;
  size_t off_;
};

// ---- records ------------------------------------------------------------------

struct Header {
  int64_t timestamp = 0;
  uint64_t entries = 0;
  uint64_t written = 0;
  Header() = default;
  explicit Header(size_t entries)
      : timestamp(std::chrono::duration_cast<std::chrono::milliseconds>(
                      std::chrono::system_clock::now().time_since_epoch())
                      .count()),
        entries(entries) {}
};

template <typename N> struct SceneMeta {
  bool suspend = false;
  bool terminate = false;
  int solverIter = 0;
  N solverStep = 0, solverScale = 0, surfaceRes = 0, gravity = 0;
  glm::tvec3<N> minBound, maxBound;
};

template <typename N> struct Scene {
  SceneMeta<N> meta;
  std::vector<fluid::Well<N>> wells;
  std::vector<fluid::Source<N>> sources;
  std::vector<fluid::Drain<N>> drains;
  std::vector<fluid::Query<N>> queries;
};

inline void put(Writer& w, const Header& h) { w.put(h.timestamp).put(h.entries).put(h.written); }
inline void get(Reader& r, Header& h) {
  h.timestamp = r.get<int64_t>();
  h.entries = r.get<uint64_t>();
  h.written = r.get<uint64_t>();
}

template <typename T, typename N> void put(Writer& w, const fluid::Particle<T, N>& p) {
  w.put(p.id).put(int32_t(p.type)).put(p.mass).put(p.colour).put3(p.position).put3(p.velocity);
}
template <typename T, typename N> void get(Reader& r, fluid::Particle<T, N>& p) {
  p.id = r.get<T>();
  p.type = fluid::Type(r.get<int32_t>());
  p.mass = r.get<N>();
  p.colour = r.get<uint32_t>();
  p.position = r.get3<N>();
  p.velocity = r.get3<N>();
}

template <typename N> void put(Writer& w, const geometry::MeshTriangle<N>& t) {
  w.put3(t.v0).put3(t.v1).put3(t.v2).put3(t.n0).put3(t.n1).put3(t.n2).put3(t.c);
}
template <typename N> void get(Reader& r, geometry::MeshTriangle<N>& t) {
  t.v0 = r.get3<N>();
  t.v1 = r.get3<N>();
  t.v2 = r.get3<N>();
  t.n0 = r.get3<N>();
  t.n1 = r.get3<N>();
  t.n2 = r.get3<N>();
  t.c = r.get3<uint32_t>();
}

template <typename N> void put(Writer& w, const fluid::QueryResult<N>& q) {
  w.put(uint64_t(q.id)).put3(q.point).put(uint64_t(q.neighbours)).put(q.averageColour);
}
template <typename N> void put(Writer& w, const fluid::Query<N>& q) {
  w.put(uint64_t(q.id)).put3(q.point);
}
template <typename N> void get(Reader& r, fluid::Query<N>& q) {
  q.id = size_t(r.get<uint64_t>());
  q.point = r.get3<N>();
}
template <typename N> void put(Writer& w, const fluid::Well<N>& x) { w.put3(x.centre).put(x.force); }
template <typename N> void get(Reader& r, fluid::Well<N>& x) {
  x.centre = r.get3<N>();
  x.force = r.get<N>();
}
template <typename N> void put(Writer& w, const fluid::Source<N>& x) {
  w.put3(x.centre).put3(x.velocity).put(uint64_t(x.rate)).put(uint64_t(x.tag)).put(x.colour);
}
template <typename N> void get(Reader& r, fluid::Source<N>& x) {
  x.centre = r.get3<N>();
  x.velocity = r.get3<N>();
  x.rate = size_t(r.get<uint64_t>());
  x.tag = size_t(r.get<uint64_t>());
  x.colour = r.get<uint32_t>();
}
template <typename N> void put(Writer& w, const fluid::Drain<N>& x) {
  w.put3(x.centre).put(x.width).put(x.depth);
}
template <typename N> void get(Reader& r, fluid::Drain<N>& x) {
  x.centre = r.get3<N>();
  x.width = r.get<N>();
  x.depth = r.get<N>();
}
template <typename N> void put(Writer& w, const geometry::Triangle<N>& t) {
  w.put3(t.v0).put3(t.v1).put3(t.v2);
}
template <typename N> void get(Reader& r, geometry::Triangle<N>& t) {
  t.v0 = r.get3<N>();
  t.v1 = r.get3<N>();
  t.v2 = r.get3<N>();
}
template <typename N> void put(Writer& w, const SceneMeta<N>& m) {
  w.put(m.suspend).put(m.terminate).put(int32_t(m.solverIter)).put(m.solverStep)
      .put(m.solverScale).put(m.surfaceRes).put(m.gravity).put3(m.minBound).put3(m.maxBound);
}
template <typename N> void get(Reader& r, SceneMeta<N>& m) {
  m.suspend = r.get<bool>();
  m.terminate = r.get<bool>();
  m.solverIter = r.get<int32_t>();
  m.solverStep = r.get<N>();
  m.solverScale = r.get<N>();
  m.surfaceRes = r.get<N>();
  m.gravity = r.get<N>();
  m.minBound = r.get3<N>();
  m.maxBound = r.get3<N>();
}

// ---- blocks (structures.hpp:391-425) ---------------------------------------------

// Header (written = 0), records, then the header again with written = entries.
template <typename Rec> size_t writeBlock(Sink sink, const std::vector<Rec>& xs, size_t offset = 0) {
  Header h(xs.size());
  Writer w(sink, offset);
  put(w, h);
  for (const Rec& x : xs) put(w, x);
  const size_t end = w.offset();
  std::atomic_thread_fence(std::memory_order_release);
  h.written = xs.size();
  Writer again(sink, offset);
  put(again, h);
  return end;
}

template <typename T, typename N>
size_t writeParticles(Sink sink, const std::vector<fluid::Particle<T, N>>& xs) {
  return writeBlock(sink, xs);
}
template <typename N>
size_t writeTriangles(Sink sink, const std::vector<geometry::MeshTriangle<N>>& xs) {
  return writeBlock(sink, xs);
}
template <typename N>
size_t writeQueries(Sink sink, const std::vector<fluid::QueryResult<N>>& xs) {
  return writeBlock(sink, xs);
}

inline std::pair<Header, size_t> readHeader(Source src, size_t offset = 0) {
  Reader r(src, offset);
  Header h;
  get(r, h);
  return {h, r.offset()};
}

template <typename Rec> std::vector<Rec> readBlock(Source src, size_t& offset) {
  Reader r(src, offset);
  Header h;
  get(r, h);
  std::vector<Rec> xs(h.entries);
  for (Rec& x : xs) get(r, x);
  offset = r.offset();
  return xs;
}

template <typename T, typename N>
std::vector<fluid::Particle<T, N>> readParticles(Source src) {
  size_t off = 0;
  return readBlock<fluid::Particle<T, N>>(src, off);
}
template <typename N> std::vector<geometry::MeshTriangle<N>> readTriangles(Source src) {
  size_t off = 0;
  return readBlock<geometry::MeshTriangle<N>>(src, off);
}

// scene.mmf: SceneMeta, then the wells, sources, drains and queries blocks.
// The reference reader sizes the query block but leaves its records unread
// (structures.hpp:467-471); they are read here.
template <typename N> void readSceneMeta(Source src, SceneMeta<N>& meta) {
  Reader r(src, 0);
  get(r, meta);
}

template <typename N> Scene<N> readScene(Source src) {
  Scene<N> s;
  Reader r(src, 0);
  get(r, s.meta);
  size_t off = r.offset();
  s.wells = readBlock<fluid::Well<N>>(src, off);
  s.sources = readBlock<fluid::Source<N>>(src, off);
  s.drains = readBlock<fluid::Drain<N>>(src, off);
  s.queries = readBlock<fluid::Query<N>>(src, off);
  return s;
}

template <typename N> size_t writeScene(Sink sink, const Scene<N>& s) {
  Writer w(sink, 0);
  put(w, s.meta);
  size_t off = w.offset();
  off = writeBlock(sink, s.wells, off);
  off = writeBlock(sink, s.sources, off);
  off = writeBlock(sink, s.drains, off);
  return writeBlock(sink, s.queries, off);
}

template <typename N> size_t writeEmptyScene(Sink sink, const SceneMeta<N>& meta) {
  Scene<N> s;
  s.meta = meta;
  return writeScene(sink, s);
}

// dynamic_bodies.mmf (structures.hpp:478-493): a header, then vec3 records.
template <typename N> fluid::RigidBody<N> readRigidBody(Source src) {
  size_t off = 0;
  Reader r(src, 0);
  Header h;
  get(r, h);
  off = r.offset();
  std::vector<glm::tvec3<N>> ts(h.entries);
  Reader rr(src, off);
  for (auto& t : ts) t = rr.get3<N>();
  return fluid::RigidBody<N>(ts);
}

// colliders.mmf (structures.hpp:495-505): a header, then triangle records.
template <typename N> fluid::MeshCollider<N> readCollider(Source src) {
  size_t off = 0;
  return fluid::MeshCollider<N>(readBlock<geometry::Triangle<N>>(src, off));
}

// ---- mapped files (miommf::padOut/createSink/createSource, structures.hpp:11-55) ----

// A file mapped shared, read-write or read-only. createSink first fills the
// file with max(4096, length) spaces, as the sample does, so a reader that
// maps it early sees a whole file.
class MappedFile {
 public:
  MappedFile() = default;
  MappedFile(const MappedFile&) = delete;
  MappedFile& operator=(const MappedFile&) = delete;
  MappedFile(MappedFile&& o) noexcept { swap(o); }
  MappedFile& operator=(MappedFile&& o) noexcept {
    swap(o);
    return *this;
  }
  ~MappedFile() {
    if (data_) ::munmap(data_, size_);
  }

  static MappedFile createSink(const std::string& path, size_t length) {
    const size_t size = length < 4096 ? 4096 : length;
    const int fd = ::open(path.c_str(), O_RDWR | O_CREAT | O_TRUNC, 0644);
    if (fd < 0) throw std::runtime_error("wire: cannot create " + path);
    std::vector<char> spaces(1 << 16, ' ');
    for (size_t done = 0; done < size;) {
      const size_t k = std::min(spaces.size(), size - done);
      const ssize_t got = ::write(fd, spaces.data(), k);
      if (got <= 0) {
        ::close(fd);
        throw std::runtime_error("wire: cannot pad " + path);
      }
      done += size_t(got);
    }
    return map(fd, size, true, path);
  }

  static MappedFile createSource(const std::string& path) {
    const int fd = ::open(path.c_str(), O_RDONLY);
    if (fd < 0) throw std::runtime_error("wire: cannot open " + path);
    struct stat st {};
    if (::fstat(fd, &st) != 0) {
      ::close(fd);
      throw std::runtime_error("wire: cannot stat " + path);
    }
    return map(fd, size_t(st.st_size), false, path);
  }

  Sink sink() { return Sink{static_cast<unsigned char*>(data_), size_}; }
  Source source() const { return Source{static_cast<const unsigned char*>(data_), size_}; }
  size_t size() const { return size_; }
  bool mapped() const { return data_ != nullptr; }

 private:
  static MappedFile map(int fd, size_t size, bool writable, const std::string& path) {
    MappedFile m;
    if (size) {
      void* p = ::mmap(nullptr, size, writable ? PROT_READ | PROT_WRITE : PROT_READ, MAP_SHARED,
                       fd, 0);
      if (p == MAP_FAILED) {
        ::close(fd);
        throw std::runtime_error("wire: cannot map " + path);
      }
      m.data_ = p;
      m.size_ = size;
    }
    ::close(fd);
    return m;
  }
  void swap(MappedFile& o) {
    std::swap(data_, o.data_);
    std::swap(size_, o.size_);
  }
  void* data_ = nullptr;
  size_t size_ = 0;
};

}  // namespace wire

#endif
