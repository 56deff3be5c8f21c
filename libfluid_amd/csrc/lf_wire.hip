// Particle and triangle blocks of libfluid's memory-mapped file wire format
// (samples/structures.hpp:391-425 writers, field order per
// structures.hpp:73-85 and 234-258), packed on the device from the context's
// sorted state and its last marching-cubes output, then copied into a host
// buffer (typically a shared mapping of particles.mmf / triangles.mmf).
//
// Records are packed little-endian with no padding: particle = id u64, type
// i32, mass f32, colour u32, position, velocity (44 B); mesh triangle = 3
// positions, 3 normals, 3 u32 colours (84 B). Both are whole 32-bit words, so
// one thread writes one word: the stores of a wave are contiguous.
#include <algorithm>
#include <atomic>
#include <cstring>

#include "lf_internal.hpp"

namespace lf {
namespace {

constexpr uint32_t kParticleWords = 11;
constexpr uint32_t kTriangleWords = 21;
constexpr unsigned kWireThreads = 256;
constexpr unsigned kWireMaxBlocks = 8192;

__global__ __launch_bounds__(kWireThreads) void k_wire_particles(
    uint64_t words, const float4* __restrict__ X, const float4* __restrict__ V,
    const uint8_t* __restrict__ T, const uint32_t* __restrict__ ID, uint32_t* __restrict__ out) {
  const uint64_t stride = uint64_t(gridDim.x) * kWireThreads;
  for (uint64_t w = uint64_t(blockIdx.x) * kWireThreads + threadIdx.x; w < words; w += stride) {
    const uint64_t i = w / kParticleWords;
    const uint32_t f = uint32_t(w - i * kParticleWords);
    const float* x = reinterpret_cast<const float*>(X + i);
    const float* v = reinterpret_cast<const float*>(V + i);
    uint32_t word;
    if (f == 0) word = ID[i];                              // id, low half
    else if (f == 1) word = 0u;                            // id, high half
    else if (f == 2) word = uint32_t(int32_t(T[i]));       // type
    else if (f == 3) word = __float_as_uint(x[3]);         // mass
    else if (f == 4) word = __float_as_uint(v[3]);         // colour (ARGB bits)
    else if (f < 8) word = __float_as_uint(x[f - 5]);      // position
    else word = __float_as_uint(v[f - 8]);                 // velocity
    out[w] = word;
  }
}

__global__ __launch_bounds__(kWireThreads) void k_wire_triangles(
    uint64_t words, const float* __restrict__ pos, const float* __restrict__ nrm,
    const uint32_t* __restrict__ col, int attributes, uint32_t* __restrict__ out) {
  const uint64_t stride = uint64_t(gridDim.x) * kWireThreads;
  for (uint64_t w = uint64_t(blockIdx.x) * kWireThreads + threadIdx.x; w < words; w += stride) {
    const uint64_t t = w / kTriangleWords;
    const uint32_t f = uint32_t(w - t * kTriangleWords);
    uint32_t word;
    if (f < 9) word = __float_as_uint(pos[9 * t + f]);
    else if (f < 18) word = attributes ? __float_as_uint(nrm[9 * t + (f - 9)]) : 0u;
    else word = attributes ? col[3 * t + (f - 18)] : 0xFFFFFFFFu;
    out[w] = word;
  }
}

void storeHeader(unsigned char* dst, int64_t timestamp, uint64_t entries, uint64_t written) {
  std::memcpy(dst, &timestamp, 8);
  std::memcpy(dst + 8, &entries, 8);
  std::memcpy(dst + 16, &written, 8);
}

// Header (written = 0), the records from the staging buffer, then the header
// with written = entries once the copy has landed.
int publish(Ctx* c, void* dst, int64_t timestamp, uint64_t entries, size_t recordBytes) {
  unsigned char* d = static_cast<unsigned char*>(dst);
  storeHeader(d, timestamp, entries, 0);
  if (recordBytes)
    LF_HIP(c, hipMemcpyAsync(d + 24, c->wireStage.p, recordBytes, hipMemcpyDeviceToHost, c->stream));
  LF_HIP(c, hipStreamSynchronize(c->stream));
  std::atomic_thread_fence(std::memory_order_release);
  storeHeader(d, timestamp, entries, entries);
  return LF_OK;
}

}  // namespace
}  // namespace lf

using namespace lf;

extern "C" {

int lf_wire_particles(lf_ctx* c, void* dst, size_t capacity, int64_t timestamp, size_t* bytes) {
  if (!c || !bytes) return LF_ERR_ARG;
  const size_t n = c->n;
  const size_t need = LF_WIRE_HEADER_BYTES + n * LF_WIRE_PARTICLE_BYTES;
  *bytes = need;
  if (!dst) return LF_OK;
  if (capacity < need)
    return fail(c, LF_ERR_CAPACITY, "lf_wire_particles: the buffer holds " +
                                        std::to_string(capacity) + " bytes, the block needs " +
                                        std::to_string(need));
  LF_HIP(c, hipSetDevice(c->device));
  const uint64_t words = uint64_t(n) * kParticleWords;
  if (n) {
    int rc = ensure(c, c->wireStage, size_t(words), "wire staging");
    if (rc) return rc;
    const int cur = c->cur;
    const unsigned grid = unsigned(std::min<uint64_t>((words + kWireThreads - 1) / kWireThreads, kWireMaxBlocks));
    rc = launch(c, "wire_particles", k_wire_particles, dim3(grid), dim3(kWireThreads), 0, words,
                (const float4*)c->X[cur].p, (const float4*)c->V[cur].p,
                (const uint8_t*)c->T[cur].p, (const uint32_t*)c->ID[cur].p, c->wireStage.p);
    if (rc) return rc;
  }
  return publish(c, dst, timestamp, n, n * LF_WIRE_PARTICLE_BYTES);
}

int lf_wire_triangles(lf_ctx* c, void* dst, size_t capacity, int64_t timestamp, size_t* bytes) {
  if (!c || !bytes) return LF_ERR_ARG;
  if (lf::mcJoin(c)) return lf::fail(c, LF_ERR_HIP, "waiting for the marching-cubes stream");
  if (!c->mcEmitted) return fail(c, LF_ERR_STATE, "lf_wire_triangles: nothing emitted");
  size_t total = 0;
  int rc = lf_mc_count(c, &total);
  if (rc) return rc;  // includes LF_ERR_CAPACITY: the device holds an incomplete mesh
  const size_t need = LF_WIRE_HEADER_BYTES + total * LF_WIRE_MESH_TRIANGLE_BYTES;
  *bytes = need;
  if (!dst) return LF_OK;
  if (capacity < need)
    return fail(c, LF_ERR_CAPACITY, "lf_wire_triangles: the buffer holds " +
                                        std::to_string(capacity) + " bytes, the block needs " +
                                        std::to_string(need));
  LF_HIP(c, hipSetDevice(c->device));
  const uint64_t words = uint64_t(total) * kTriangleWords;
  if (total) {
    if ((rc = ensure(c, c->wireStage, size_t(words), "wire staging"))) return rc;
    const int attr = (c->triNrm.p != nullptr && c->mcAttr) ? 1 : 0;
    const unsigned grid = unsigned(std::min<uint64_t>((words + kWireThreads - 1) / kWireThreads, kWireMaxBlocks));
    rc = launch(c, "wire_triangles", k_wire_triangles, dim3(grid), dim3(kWireThreads), 0, words,
                (const float*)c->triPos.p, (const float*)c->triNrm.p,
                (const uint32_t*)c->triCol.p, attr, c->wireStage.p);
    if (rc) return rc;
  }
  return publish(c, dst, timestamp, total, total * LF_WIRE_MESH_TRIANGLE_BYTES);
}

}  // extern "C"
