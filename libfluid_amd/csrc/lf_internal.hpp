// Internal declarations of the gfx950 PBF + marching-cubes library: the
// context, device-side math shared by the kernels, and the host-side launchers
// each .hip file provides. Paths cited are relative to the libfluid reference
// root (tom91136/libfluid).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "libfluid_amd.h"

// ---------------------------------------------------------------------------
// Arithmetic contract (identical op order in every kernel so the FP32 path is
// reproducible bit for bit against the CPU restatement):
//  * the library is compiled with -ffp-contract=off; each fused multiply-add
//    is an explicit __builtin_fmaf;
//  * sqrt and division are IEEE correctly rounded (hipcc's default
//    -fhip-fp32-correctly-rounded-divide-sqrt; the pair terms use shortened
//    forms of the same sequences where they give the same bits, kArithRn);
//  * d2 = fma(dz,dz, fma(dy,dy, dx*dx)) with d = p_a - p_b;
//  * W6(r) = poly6 * t*t*t, t = fma(-r, r, h*h), for r <= h
//    gradW  = d * (spiky * ((h-r)*(h-r) / r)) for 1e-8 <= r <= h
//    (oclsph_kernel.h:50-71);
//  * "r <= h" is tested on d2 against rmax2, the largest float whose
//    correctly rounded square root is <= h, which is the same predicate
//    without the square root;
//  * divisions by constants are multiplications by the reciprocal folded on
//    the host (W6(r) / W6(dq) and / rho0 of sph_delta);
//  * XSPH takes W6 on d2 (t = h*h - d2, no sqrt); the density field takes
//    t = fma(-dx,dx, fma(-dy,dy, fma(-dz,dz, h*h))) and rho = fma((m poly6) t, t t, rho).
// The FAST variants of the pair kernels (LF_FLAG_FAST_MATH) replace the
// correctly rounded sqrt and 1/r by the hardware approximations (1 ulp);
// they are checked against the oracle with a tolerance instead.
// ---------------------------------------------------------------------------

namespace lf {

// Constants folded on the host from lf_pbf_params (FP32, the reference's
// folding order of oclsph_kernel.h:20-51).
struct PbfConsts {
  float h, hh, rmax2, poly6, spiky, rhoRecip, rho0, cfm, corrK, wdq, wdqRecip, invDt, dt, scale,
      vd, xsph, xsphRho;  // xsphRho = xsph / rho0
  float listR2;      // neighbour-list radius^2, (h + skin)^2 with rounding margin
  float skinSafe;    // skin less a rounding margin
  float listR2R;     // the same for lists rebuilt by a lambda pass (a wider skin)
  float skinSafeR;
  float minExt[3];   // minBound/scale - 2h (oclsph.hpp:359)
  float minB[3], maxB[3];
  float grav[3];
  uint32_t G;        // dense cell-table size Morton(extent) (oclsph.hpp:749)
  uint32_t extent[3];
  uint32_t flags;
  // this scene's slice of the concatenated cell tables of a batch (0 for a
  // single scene). The sort key is keyBase + min(key, G): keys >= G belong to
  // no cell and share the scene's last bucket, so scenes sort apart and each
  // keeps the reference's order of its cells.
  uint32_t keyBase;
  uint32_t rnDiv;    // 1 when 1e-4 <= h <= 1e4 (the kArithRn forms are exact)
  float densityTol;  // > 0: skip the remaining passes once the density error is below it
  float qUnit;       // LF_FLAG_PACKED16: one fixed-point quantum, h / 65536 (solver units)
  double qInv;       // 65536 / h
};

// ---- LF_FLAG_PACKED16: cell-relative 16-bit fixed-point positions ---------------
// The lambda and delta passes gather their neighbours' p* as one 8-byte
// record instead of a 16-byte float4: per axis the 21-bit fixed-point
// coordinate round((p - minExt) * 65536 / h) mod 2^21, i.e. the 16-bit offset
// of p within its cell (unit h / 65536) and the low 5 bits of its cell
// coordinate. The difference of two such coordinates, taken mod 2^21 and
// sign-extended, is p_a - p_b in quanta exactly (pairs closer than 16 cells on
// every axis: candidates are within one cell at the list build), so each
// relative position is the true one to one quantum per axis, a finer grain
// than FP32 absolute coordinates far from the origin.
constexpr uint32_t kQBits = 21;
constexpr uint32_t kQMask = (1u << kQBits) - 1u;
__device__ inline uint32_t quantAxis(const PbfConsts& c, float p, int k) {
  const double q = rint((double(p) - double(c.minExt[k])) * c.qInv);
  // only positions far outside the grid leave the int32 range: clamped
  const double qc = fmin(fmax(q, -1073741824.0), 1073741823.0);
  return uint32_t(int32_t(qc)) & kQMask;
}
__device__ inline uint2 packPos(const PbfConsts& c, float x, float y, float z) {
  const uint32_t qx = quantAxis(c, x, 0), qy = quantAxis(c, y, 1), qz = quantAxis(c, z, 2);
  return make_uint2(qx | (qy << kQBits), (qy >> (32 - kQBits)) | (qz << (2 * kQBits - 32)));
}
struct QPos {
  uint32_t x, y, z;
};
__device__ inline QPos unpackPos(uint2 r) {
  return QPos{r.x & kQMask, ((r.x >> kQBits) | (r.y << (32 - kQBits))) & kQMask,
              (r.y >> (2 * kQBits - 32)) & kQMask};
}
// a - b of two fixed-point coordinates, in solver units
__device__ inline float qDiff(uint32_t a, uint32_t b, float unit) {
  const int32_t d = int32_t((a - b) << (32 - kQBits)) >> (32 - kQBits);
  return float(d) * unit;
}

// ---- work decomposition -------------------------------------------------------
// Chunks of at most the workgroup size of consecutive particles of one scene
// (sorted and unsorted order keep each scene's range), so the constants a
// workgroup reads are uniform.
struct Chunk {
  uint32_t scene, begin, end, index;  // index: position in the chunk list
};
struct Work {
  const PbfConsts* consts;  // per scene
  const Chunk* chunks;
  uint32_t nChunks;
  uint32_t* queue;  // persistent launch: [0] next chunk, [1] blocks done; null: a workgroup per chunk
  const uint32_t* skip;  // per scene: non-zero skips the scene's chunks (converged); may be null
  // non-null: the health counters of the step (non-finite p*, overfull
  // cells); the launch does nothing once either is raised (a diverged state)
  const uint32_t* halt;
  uint32_t ppt;  // chunks per workgroup of a launch without a queue (particles per thread)
};

// Workgroups are dispatched to the 8 XCDs round robin (workgroup b to XCD
// b % 8), each XCD with its own L2. This maps them so that XCD x works through
// stripes of LF_XCD_STRIPE consecutive chunks (4,096 particles at 256 per
// chunk), stripe j on XCD j % 8: neighbouring particles, whose data the pair
// passes gather, mostly share an L2, and a region of costly chunks (a splash
// whose lists are rebuilt) is spread over every XCD. Contiguous eighths per
// XCD (LF_XCD_STRIPE 0) left that work to one XCD: 1.936 against 1.783
// ms/step over steps 5-25 of the 1M dam break (stripes of 4: 1.782, of 1, the
// identity: 1.794). A bijection on [0, n).
#ifndef LF_XCD_STRIPE
#define LF_XCD_STRIPE 16
#endif
__device__ inline uint32_t xcdChunk(uint32_t b, uint32_t n) {
  const uint32_t x = b & 7u, i = b >> 3;
  if constexpr (LF_XCD_STRIPE == 0) {
    const uint32_t q = n >> 3, r = n & 7u;
    return x * q + min(x, r) + i;
  } else {
    // stripes of S chunks dealt to the XCDs in turn (S = 1: the identity)
    constexpr uint32_t S = LF_XCD_STRIPE;
    if (b >= (n / (8u * S)) * (8u * S)) return b;
    return (i / S) * (8u * S) + x * S + i % S;
  }
}

// Runs body(consts, chunk) for the chunks of this workgroup: one chunk per
// workgroup, or, persistent, chunks taken from the shared counter until it
// passes nChunks (every workgroup reaches that exit). The last workgroup to
// leave resets the counters for the next launch, so no host-side clear is
// needed (graph-safe).
template <typename Body> __device__ inline void forEachChunk(const Work& W, Body&& body) {
  if (W.halt && (W.halt[0] | W.halt[1])) return;  // uniform: before any chunk is taken
  // The constants are copied out of global memory once per chunk: a reference
  // into global memory would be re-read after every store the body makes.
  if (!W.queue) {
    // workgroup b takes chunks ppt*x .. ppt*x + ppt - 1, x = xcdChunk(b): one
    // after another (each thread its particle of each), so a thread handles
    // up to ppt particles; the trip count is uniform over the workgroup
    const uint32_t nGroups = (W.nChunks + W.ppt - 1u) / W.ppt;
    const uint32_t first = xcdChunk(blockIdx.x, nGroups) * W.ppt;
    for (uint32_t k = 0; k < W.ppt; ++k) {
      if (first + k >= W.nChunks) break;
      const Chunk ch = W.chunks[first + k];
      if (W.skip && W.skip[ch.scene]) continue;
      const PbfConsts c = W.consts[ch.scene];
      body(c, ch);
    }
    return;
  }
  __shared__ uint32_t next;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) next = atomicAdd(&W.queue[0], 1u);
    __syncthreads();
    const uint32_t i = next;
    if (i >= W.nChunks) break;
    const Chunk ch = W.chunks[i];
    if (W.skip && W.skip[ch.scene]) continue;
    const PbfConsts c = W.consts[ch.scene];
    body(c, ch);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&W.queue[1], 1u) == gridDim.x - 1) {
      W.queue[0] = 0u;
      W.queue[1] = 0u;
      __threadfence();
    }
  }
}

// Counter pairs of the persistent launches (one per kernel kind).
enum { kSlotList = 0, kSlotLambda = 1, kSlotDelta = 2, kSlotXsph = 3, kQueueSlots = 4 };

int foldConsts(const lf_pbf_params& p, float skinFrac, PbfConsts& c, std::string& err);

// Neighbour-list capacity per particle (~50 within h + 0.2h at rest with
// h = 2 spacings). A particle with more keeps the count and enumerates its 27
// cells instead; results are the same either way.
constexpr uint32_t kMaxNbr = 128;

// List storage. Entry i of particle a holds its neighbour b as the 24-bit
// value b - a + 2^23 (neighbours more than 2^23 sorted positions away mark the
// list overflowed, so a context of up to 2^23 + 1 particles never needs it).
// Entries come in groups of 8 (24 bytes, two 12-byte loads per lane), the 64
// particles of a wave side by side within a group (1,536 contiguous bytes),
// groups in entry order. Entries are 3/4 the size of 32-bit indices: the pair
// passes stream the lists, most of their memory traffic.
constexpr uint32_t kNbrGroup = 8;
constexpr uint32_t kNbrGroupBytes = kNbrGroup * 3;           // 24
constexpr uint32_t kNbrWaveGroupBytes = 64 * kNbrGroupBytes;  // 1536
constexpr uint32_t kNbrBias = 1u << 23;
// Byte offset of group g (entries 8g .. 8g+7) of particle a's list.
__host__ __device__ constexpr uint64_t nbrGroupByte(uint32_t a, uint32_t g) {
  return (uint64_t(a >> 6) * (kMaxNbr / kNbrGroup) + g) * kNbrWaveGroupBytes +
         uint64_t(a & 63u) * kNbrGroupBytes;
}
// Capacity in 32-bit words.
inline size_t nbrCapacity(size_t n) {
  return ((n + 63) / 64) * (kMaxNbr / kNbrGroup) * kNbrWaveGroupBytes / 4;
}
// Packing of 8 biased entries into 6 words and back.
__device__ inline void packGroup(const uint32_t e[kNbrGroup], uint32_t w[6]) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t* x = e + 4 * h;
    w[3 * h + 0] = x[0] | (x[1] << 24);
    w[3 * h + 1] = (x[1] >> 8) | (x[2] << 16);
    w[3 * h + 2] = (x[2] >> 16) | (x[3] << 8);
  }
}
__device__ inline void unpackGroup(const uint32_t w[6], uint32_t e[kNbrGroup]) {
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t* x = w + 3 * h;
    e[4 * h + 0] = x[0] & 0xFFFFFFu;
    e[4 * h + 1] = ((x[0] >> 24) | (x[1] << 8)) & 0xFFFFFFu;
    e[4 * h + 2] = ((x[1] >> 16) | (x[2] << 16)) & 0xFFFFFFu;
    e[4 * h + 3] = x[2] >> 8;
  }
}

// ---- Morton codes (curves.h:56-72), 10 bits per axis ----------------------
__host__ __device__ inline uint32_t spread10(uint32_t v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__host__ __device__ inline uint32_t mortonEncode(uint32_t x, uint32_t y, uint32_t z) {
  return spread10(x) | (spread10(y) << 1) | (spread10(z) << 2);
}
constexpr uint32_t kMortonX = 0x09249249u;  // x bits of a 30-bit code
constexpr uint32_t kMortonY = kMortonX << 1;
constexpr uint32_t kMortonZ = kMortonX << 2;

// Dilated-integer step of one axis inside a Morton code: the axis bits of
// code+/-1 along that axis, wrapping mod 1024 like the reference's size_t
// arithmetic masked to 10 bits (curves.h:56-72).
__device__ inline uint32_t dilatedInc(uint32_t code, uint32_t mask) {
  return ((code | ~mask) + 1u) & mask;
}
__device__ inline uint32_t dilatedDec(uint32_t code, uint32_t mask) {
  return ((code & mask) - 1u) & mask;
}

// static_cast<size_t>((p - minExt)/h) of oclsph.hpp:327-332 masked to 10 bits;
// negative or NaN offsets (undefined in the reference) map to 0.
__device__ inline uint32_t cellCoord(float q) {
  if (!(q >= 0.0f)) return 0u;
  if (q >= 4294967296.0f) return 1023u;
  return static_cast<uint32_t>(q) & 1023u;
}

__device__ inline uint32_t cellKey(const PbfConsts& c, float px, float py, float pz) {
  const uint32_t qx = cellCoord((px - c.minExt[0]) / c.h);
  const uint32_t qy = cellCoord((py - c.minExt[1]) / c.h);
  const uint32_t qz = cellCoord((pz - c.minExt[2]) / c.h);
  return mortonEncode(qx, qy, qz);
}

__device__ inline float dist2(float dx, float dy, float dz) {
  return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

__device__ inline float poly6W(const PbfConsts& c, float r) {
  const float t = __builtin_fmaf(-r, r, c.hh);
  return c.poly6 * ((t * t) * t);
}
// Arithmetic of the pair terms' square root and division (a template
// argument of the lambda and delta passes):
//  * kArithIeee: hipCC's IEEE sequences (sqrtf, FP32 division), for any h;
//  * kArithRn: sqrtRn and divRn below, the same bits in fewer instructions
//    when 1e-4 <= h <= 1e4 (PbfConsts::rnDiv);
//  * kArithHw: the hardware approximations (LF_FLAG_FAST_MATH).
enum Arith : int { kArithIeee = 0, kArithRn = 1, kArithHw = 2 };

// The kArithRn forms are the compiler's own sequences with the steps that only
// act near the ends of the FP32 range left out:
//  * sqrtRn: for x >= 2^-96 the compiler's sqrtf takes the hardware estimate
//    s (within 1 ulp) and returns s - 1 ulp, s or s + 1 ulp by the signs of
//    the residuals x - s' * s; sqrtRn does exactly that. Below 2^-96 sqrtf
//    first scales x by 2^32 and passes 0 through; sqrtRn does not, and returns
//    some r <= 2^-47 where sqrtf gives one below 2^-48. For a pair term the
//    two agree: r < 1e-8 gives a zero spiky term either way, and r * r is far
//    below half an ulp of h * h >= 1e-8, so fma(-r, r, h * h) = h * h.
//  * divRn: the compiler's FP32 division refines the hardware reciprocal y
//    of b once (y += y * (1 - b y)), takes q = a y and corrects it twice
//    (q += y * (a - b q)). Its scale steps (v_div_scale, v_div_fmas) act only
//    when b or 1/b is denormal, a is below 2^-103, or a/b is denormal or 2^96
//    beyond in exponent, and its fix-up step only settles zeros, infinities
//    and NaNs. For the spiky term's a = (h - r)^2, b = r, 1e-8 <= r <= h and
//    1e-4 <= h <= 1e4, a is 0 or at least (h 2^-25)^2 and none of these
//    occurs; a = 0 gives +0 either way.
__device__ inline float sqrtRn(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sd = __uint_as_float(__float_as_uint(s) - 1u);
  const float su = __uint_as_float(__float_as_uint(s) + 1u);
  const float rd = __builtin_fmaf(-sd, s, x);
  const float ru = __builtin_fmaf(-su, s, x);
  const float t = rd <= 0.0f ? sd : s;
  return ru > 0.0f ? su : t;
}
__device__ inline float divRn(float a, float b) {
  float y = __builtin_amdgcn_rcpf(b);
  y = __builtin_fmaf(__builtin_fmaf(-b, y, 1.0f), y, y);
  float q = a * y;
  q = __builtin_fmaf(__builtin_fmaf(-b, q, a), y, q);
  return __builtin_fmaf(__builtin_fmaf(-b, q, a), y, q);
}

// caller guarantees r <= h
__device__ inline float spikyScale(const PbfConsts& c, float r) {
  if (!(r >= 1e-8f)) return 0.0f;
  const float q = c.h - r;
  return c.spiky * ((q * q) / r);
}

template <int AR> __device__ inline float sqrtM(float x) {
  return AR == kArithHw ? __builtin_amdgcn_sqrtf(x) : AR == kArithRn ? sqrtRn(x) : __builtin_sqrtf(x);
}
// (the shortened forms compute the quotient for every r and select: no branch)
template <int AR> __device__ inline float spikyScaleM(const PbfConsts& c, float r) {
  if (AR == kArithIeee) return spikyScale(c, r);
  const float q = c.h - r;
  const float v = c.spiky * (AR == kArithRn ? divRn(q * q, r) : (q * q) * __builtin_amdgcn_rcpf(r));
  return r >= 1e-8f ? v : 0.0f;
}

// 27-cell visit in the reference order (oclsph_kernel.h:150-192): x fastest,
// then y, then z; a cell c holds [table[c], table[c+1]) when c+1 < G and is
// empty otherwise (the reference's end = start, :186-187).
template <typename Fn>
__device__ inline void forEachNeighbourCell(uint32_t key, const uint32_t* __restrict__ table,
                                            uint32_t G, Fn&& fn) {
  const uint32_t xs[3] = {dilatedDec(key, kMortonX), key & kMortonX, dilatedInc(key, kMortonX)};
  const uint32_t ys[3] = {dilatedDec(key, kMortonY), key & kMortonY, dilatedInc(key, kMortonY)};
  const uint32_t zs[3] = {dilatedDec(key, kMortonZ), key & kMortonZ, dilatedInc(key, kMortonZ)};
#pragma unroll
  for (int iz = 0; iz < 3; ++iz) {
#pragma unroll
    for (int iy = 0; iy < 3; ++iy) {
      const uint32_t yz = ys[iy] | zs[iz];
#pragma unroll
      for (int ix = 0; ix < 3; ++ix) {
        const uint32_t code = xs[ix] | yz;
        if (code + 1u < G) fn(table[code], table[code + 1u]);
      }
    }
  }
}

// The same 27 cells in the same order with the cell loop kept rolled (the
// body is emitted once instead of 27 times, for callers whose body is large
// and runs for every cell: the list build), and the next cell's table
// entries loaded before the current cell is visited, so their latency
// overlaps the visit.
template <typename Fn>
__device__ inline void forEachNeighbourCellRolled(uint32_t key, const uint32_t* __restrict__ table,
                                                  uint32_t G, Fn&& fn) {
  const uint32_t xs[3] = {dilatedDec(key, kMortonX), key & kMortonX, dilatedInc(key, kMortonX)};
  const uint32_t ys[3] = {dilatedDec(key, kMortonY), key & kMortonY, dilatedInc(key, kMortonY)};
  const uint32_t zs[3] = {dilatedDec(key, kMortonZ), key & kMortonZ, dilatedInc(key, kMortonZ)};
  auto codeOf = [&](uint32_t i) {
    const uint32_t ix = i % 3u, iy = (i / 3u) % 3u, iz = i / 9u;
    return (ix == 0 ? xs[0] : (ix == 1 ? xs[1] : xs[2])) |
           (iy == 0 ? ys[0] : (iy == 1 ? ys[1] : ys[2])) |
           (iz == 0 ? zs[0] : (iz == 1 ? zs[1] : zs[2]));
  };
  uint32_t code = codeOf(0);
  bool valid = code + 1u < G;
  uint32_t s = valid ? table[code] : 0u, e = valid ? table[code + 1u] : 0u;
#pragma unroll 1
  for (uint32_t i = 0; i < 27u; ++i) {
    uint32_t sN = 0u, eN = 0u;
    bool validN = false;
    if (i + 1u < 27u) {
      const uint32_t codeN = codeOf(i + 1u);
      validN = codeN + 1u < G;
      if (validN) {
        sN = table[codeN];
        eN = table[codeN + 1u];
      }
    }
    if (valid) fn(s, e);
    s = sN;
    e = eN;
    valid = validN;
  }
}

// p* clamped to the bounds as the delta pass clamps it (oclsph_kernel.h:364-374).
__device__ inline float clampToBounds(const PbfConsts& c, int k, float p) {
  float pos = p * c.scale;
  pos = (c.minB[k] < pos) ? pos : c.minB[k];
  pos = (pos < c.maxB[k]) ? pos : c.maxB[k];
  return pos / c.scale;
}

// ---- predict (advectAndCopy, oclsph.hpp:289-325) ---------------------------
__device__ inline void predictPstar(const PbfConsts& c, float4 x, float4 v, uint8_t type,
                                    const lf_well* __restrict__ wells, int nw, float& px,
                                    float& py, float& pz) {
  if (type == LF_OBSTACLE) {
    px = x.x / c.scale;
    py = x.y / c.scale;
    pz = x.z / c.scale;
    return;
  }
  const float m = x.w;
  float f[3] = {m * c.grav[0], m * c.grav[1], m * c.grav[2]};
  for (int i = 0; i < nw; ++i) {  // oclsph.hpp:307-317
    const float d[3] = {wells[i].centre[0] - x.x, wells[i].centre[1] - x.y,
                        wells[i].centre[2] - x.z};
    const float dist = __builtin_sqrtf(dist2(d[0], d[1], d[2]));
    if (dist < 75.0f) {
      const float inv = 1.0f / dist;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        float fk = ((d[k] * inv) * wells[i].force) * m;
        fk = fk / (dist * dist);
        fk = fminf(fmaxf(fk, -10.0f), 10.0f);
        f[k] += fk;
      }
    }
  }
  const float vx = __builtin_fmaf(f[0], c.dt, v.x);
  const float vy = __builtin_fmaf(f[1], c.dt, v.y);
  const float vz = __builtin_fmaf(f[2], c.dt, v.z);
  px = __builtin_fmaf(vx, c.dt, x.x / c.scale);
  py = __builtin_fmaf(vy, c.dt, x.y / c.scale);
  pz = __builtin_fmaf(vz, c.dt, x.z / c.scale);
}

// ---- mesh colliders (oclsph_collision.h) --------------------------------------
// Per-triangle values of collideTriangle2 / inTrig that do not depend on the
// particle, folded once on the host in FP32 with the same operations as the
// CPU restatement: nn = n / |n| with n = (b - a) x (c - a), dn = nn . a,
// v0 = c - a, v1 = b - a, dot00, dot01, dot11, invDenom = 1 / (dot00 dot11 -
// dot01^2); dot3(u, v) = fma(u.z, v.z, fma(u.y, v.y, u.x * v.x)).
struct ColliderTri {
  float a[3], nn[3], v0[3], v1[3];
  float dn, dot00, dot01, dot11, invDenom;
  float pad[3];
};

__host__ __device__ inline float dot3(const float* u, const float* v) {
  return __builtin_fmaf(u[2], v[2], __builtin_fmaf(u[1], v[1], u[0] * v[0]));
}

// collideTriangle2 (oclsph_collision.h:25-58) for one triangle: when its
// plane lies within 40 world units of p (t = dn - nn.p) and the projection
// p0 = p + nn t falls inside it (inTrig, :7-22: u >= 0, v >= 0, u + v < 1), p
// moves to p0 and v reflects about the plane (v - nn (2 v.nn)).
__device__ inline bool tryTriangle(const ColliderTri& T, float p[3], float v[3]) {
  const float t = T.dn - dot3(T.nn, p);
  if (!(-40.0f <= t && t <= 40.0f)) return false;
  float p0[3], v2[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p0[k] = __builtin_fmaf(T.nn[k], t, p[k]);
    v2[k] = p0[k] - T.a[k];
  }
  const float dot02 = dot3(T.v0, v2), dot12 = dot3(T.v1, v2);
  const float u = (T.dot11 * dot02 - T.dot01 * dot12) * T.invDenom;
  const float w = (T.dot00 * dot12 - T.dot01 * dot02) * T.invDenom;
  if (!(u >= 0.0f && w >= 0.0f && u + w < 1.0f)) return false;
  const float s = 2.0f * dot3(v, T.nn);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    p[k] = p0[k];
    v[k] = __builtin_fmaf(-T.nn[k], s, v[k]);
  }
  return true;
}

// The colliders of a delta pass. The first triangle in list order that hits
// wins (oclsph_collision.h:45-57). With a grid, the point's cell lists, in
// ascending order, every triangle that can hit anywhere in the cell (a
// conservative superset), so the first hit found there is the first hit of
// the whole list; points outside the grid scan the whole list.
struct ColliderGrid {
  const ColliderTri* tris;
  const uint32_t* cellStart;  // cells + 1 offsets into cellTris
  const uint32_t* cellTris;   // triangle indices, ascending within a cell
  float origin[3];            // world units
  float invCell;
  uint32_t dims[3];
  uint32_t n;                 // triangles (0: no colliders)
  uint32_t useGrid;           // 0: scan the whole list
};

__device__ inline bool collide(const ColliderGrid& g, float p[3], float v[3]) {
  if (g.useGrid) {
    const float qx = (p[0] - g.origin[0]) * g.invCell;
    const float qy = (p[1] - g.origin[1]) * g.invCell;
    const float qz = (p[2] - g.origin[2]) * g.invCell;
    if (qx >= 0.0f && qx < float(g.dims[0]) && qy >= 0.0f && qy < float(g.dims[1]) &&
        qz >= 0.0f && qz < float(g.dims[2])) {
      const uint32_t cell =
          (uint32_t(qz) * g.dims[1] + uint32_t(qy)) * g.dims[0] + uint32_t(qx);
      const uint32_t e = g.cellStart[cell + 1u];
      uint32_t j = g.cellStart[cell];
      if (j >= e) return false;
      // the next triangle's loads are issued before the current one is tested
      ColliderTri next = g.tris[g.cellTris[j]];
      while (true) {
        const ColliderTri cur = next;
        const bool more = ++j < e;
        if (more) next = g.tris[g.cellTris[j]];
        if (tryTriangle(cur, p, v)) return true;
        if (!more) return false;
      }
    }
  }
  for (uint32_t i = 0; i < g.n; ++i)
    if (tryTriangle(g.tris[i], p, v)) return true;
  return false;
}

// ---- context ----------------------------------------------------------------
template <typename T> struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;  // elements
};

struct KernelStat {
  uint64_t count = 0;
  double ms = 0.0;
};

struct PendingTiming {
  const char* name;
  hipEvent_t start, stop;
};

constexpr uint32_t kMcTiles = 4;  // density-field tile shapes (lf_mc.hip)

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t launchStream = nullptr;  // non-null: launch() targets this stream
  // The density-field marching cubes (sampling, classification, scan,
  // emission) run on mcStream, created on first use, after an event at the
  // end of the step, beside the next step, which touches none of their
  // buffers (P and the table alternate between steps, see pairGuard).
  // mcBusy: mcStream has work the main stream has not yet waited for
  // (mcJoin); every reader of the mesh outputs and every writer of the field
  // on the main stream joins first.
  hipStream_t mcStream = nullptr;
  hipEvent_t evMcIn = nullptr, evMcDone = nullptr;
  bool mcBusy = false;
  // evPosFinal: recorded on the main stream after every kernel that writes P
  // or the table (grid build, delta, finalise), so a density field waits for
  // the last of them rather than for everything issued since (XSPH);
  // posEventValid: no other write of P or the table came after it.
  // fieldOnMain: the field was last written on the main stream (an upload or
  // the metaball sampling), so classification waits for the main stream.
  hipEvent_t evPosFinal = nullptr;
  bool posEventValid = false;
  bool fieldOnMain = false;
  std::string err;

  size_t n = 0;
  int cur = 0;  // which state buffer holds the current state
  DevBuf<float4> X[2], V[2];  // world position + mass; velocity + colour bits
  DevBuf<uint8_t> T[2];       // particle type
  DevBuf<uint32_t> ID[2];     // upload index
  DevBuf<uint32_t> sortKey[2], sortVal[2];
  DevBuf<float4> P, PL;       // p* (+mass), p* + lambda
  // LF_FLAG_PACKED16: p* as 8-byte fixed-point records (Q: current, QL: the
  // copy the lambda pass makes for the delta pass) and lambda alone
  DevBuf<uint2> Q, QL;
  DevBuf<float> LAM;
  bool gridPacked = false;    // the last grid build wrote Q
  DevBuf<float> rho;          // density of the last lambda pass
  DevBuf<uint32_t> nbr;       // neighbour lists (nbrGroupByte layout)
  DevBuf<uint32_t> nbrCount;  // list length per particle (> kMaxNbr: overflowed)
  DevBuf<float4> P0;          // p* at the list build
  DevBuf<uint8_t> listEpoch;  // per particle: the pass its list was (re)built from
  DevBuf<float4> listR;       // per particle: p* of that pass (epoch > 1)
  int passIdx = 1;            // lambda/delta pass of the step (1-based), XSPH: iterations + 1
  DevBuf<uint32_t> maxDisp;   // [0]: float bits of the largest displacement since
  DevBuf<uint32_t> cellDisp;  // per cell: float bits of its particles' largest one
  DevBuf<uint32_t> diffused;
  DevBuf<uint32_t> table;
  // P and the table alternate between steps (gridBuild swaps them): a density
  // field of the last step may still read the other pair on mcStream. Each
  // pair carries the event recorded after the last such read (swapped with it).
  DevBuf<float4> Palt;
  DevBuf<uint32_t> tableAlt;
  struct PairGuard {
    hipEvent_t ev = nullptr;
    bool pending = false;
  } pairGuard, pairGuardAlt;
  // grid build (lf_grid.hip): bucket counts (all zero between steps), scan
  // tile sums, the list of buckets too large for the in-place ordering
  DevBuf<uint32_t> cellCount, bucketTile, bigBuckets;
  // health of the state: [0] particles whose p* is not finite, [1] cells of
  // the table holding more than guardCell particles, [2] the largest such
  // count; accumulated since the last upload. Once [0] or [1] is raised the
  // pair passes and the density field do nothing and the step calls report
  // LF_ERR_DIVERGED. healthHost: a pinned host copy each grid build writes.
  DevBuf<uint32_t> health;
  uint32_t guardCell = 256;
  uint32_t* healthHost = nullptr;
  DevBuf<lf_well> wells;
  int nWells = 0;
  DevBuf<ColliderTri> colliders;
  uint32_t nColliders = 0;
  std::vector<lf_triangle> hostColliders;   // as given, for the grid's inclusion tests
  DevBuf<double> colliderBounds;            // per triangle: inclusion-test record
  DevBuf<uint32_t> colliderCellStart, colliderCellTris, colliderCounts;
  ColliderGrid colliderGrid{};              // device pointers filled in when built
  bool colliderGridValid = false;
  float colliderGridBox[6] = {0, 0, 0, 0, 0, 0};  // scene bounds the grid was built for
  std::vector<lf_source> sources;
  std::vector<lf_drain> hostDrains;
  DevBuf<lf_drain> drains;
  DevBuf<float> queryPts;       // 3 per query, world units
  DevBuf<uint32_t> queryOut;    // per query: count, packed colour
  uint32_t nQueries = 0;
  bool queriesAnswered = false;
  uint32_t nextId = 0;          // upload index of the next appended particle
  DevBuf<uint32_t> compactBlock;  // per-block counts -> offsets of the drain compaction
  DevBuf<uint32_t> wireStage;     // packed wire records before the copy to the host
  std::vector<lf_well> hostWells;

  bool gridValid = false;
  bool stepDone = false;
  float skinFrac = 0.15f;  // neighbour-list skin / h
  bool listFresh = false;  // P still holds the predicted p* the lists were built from
  PbfConsts consts{};      // scene 0 (single-scene calls), or the scene of lf_mc_* calls
  lf_pbf_params params{};
  uint32_t gridG = 0;      // table size of the last grid build

  // scenes (a single-scene context is a batch of one)
  bool batch = false;                      // scenes and params fixed by lf_batch_upload
  std::vector<lf_pbf_params> sceneParams;  // per scene
  std::vector<PbfConsts> sceneConsts;      // per scene, folded
  std::vector<uint32_t> sceneBegin;        // particle offsets, nScenes + 1 entries
  uint32_t tableTotal = 0;                 // concatenated table size (LF_ARR_TABLE)
  uint32_t bucketTotal = 0;                // sort buckets: every scene's G cells + 1
  DevBuf<PbfConsts> dConsts;
  DevBuf<Chunk> chunks;
  uint32_t nChunks = 0;
  uint32_t chunkThreads = 256;             // workgroup size of the particle kernels
  uint32_t chunkPpt = 1;                   // chunks per workgroup (particles per thread)
  uint32_t mcTile = 0;                     // density-field tile shape (lf_exec_params), < kMcTiles
  lf_mc_params lastDensity{};              // the last density-field lattice sampled (autotuner)
  bool haveDensity = false;
  std::vector<PbfConsts> uploadedConsts;  // what dConsts holds
  std::vector<uint32_t> uploadedBegin;     // scene offsets of the uploaded chunk list
  uint32_t chunksBuiltFor = 0;             // chunkThreads the chunk list was built with
  bool persistent = false;                 // work-queue launches for the pair kernels
  uint32_t persistentBlocks = 2048;        // grid of a persistent launch
  DevBuf<uint32_t> queues;                 // 2 counters per persistent launch slot
  uint32_t gridTotal = 0;                  // tableTotal of the last grid build
  std::vector<uint32_t> chunkOffHost;      // first chunk of each scene, nScenes + 1 entries
  DevBuf<uint32_t> chunkOff;
  DevBuf<double> chunkSums;                // kSumWidth doubles per chunk
  DevBuf<double> sceneStats;               // kStatSlots x kSumWidth doubles per scene
  DevBuf<uint32_t> sceneDone;              // per scene: converged in this step
  DevBuf<uint32_t> scenePasses;            // per scene: lambda passes run in the last step
  bool converge = false;                   // some scene has density_tol > 0
  bool pairRn = true;                      // every scene has rnDiv: kArithRn pair terms
  float mcScale = 1.0f;                    // world scale of the marching-cubes output
  DevBuf<uint32_t> localKey;               // scene-local Morton key per sorted particle
  int mcScene = 0;
  bool capturing = false;
  hipGraphExec_t graph = nullptr;
  int graphSteps = 0;
  int keysIdx = 0;  // which sortKey/sortVal holds the sorted keys and order

  // marching cubes
  DevBuf<float> field;
  DevBuf<float4> mcNormal;
  DevBuf<uint32_t> mcColour;
  DevBuf<uint8_t> mcCases;
  DevBuf<uint32_t> mcBlock;   // per-block triangle counts -> offsets
  DevBuf<uint32_t> mcTotal;   // [0] = total triangles of the last emit
  DevBuf<float> triPos;       // 9 floats per triangle
  DevBuf<float> triNrm;       // 9 floats per triangle (attributes)
  DevBuf<uint32_t> triCol;    // 3 per triangle (attributes)
  size_t triCap = size_t(16) << 20;
  uint32_t S[3] = {0, 0, 0};
  float mcOrigin[3] = {0, 0, 0}, mcStep[3] = {0, 0, 0};
  bool mcSampled = false, mcCounted = false, mcEmitted = false;
  int mcAttr = 0;
  bool keepCases = false;
  uint32_t* mcScratch = nullptr;  // lf_mc_mesh_static's block counts on this device
  // per tile of the last density field: 1 = every sample of the tile was
  // written as zero (its halo holds no particle); valid while the field is
  // that density field (mcTileFlags), with the tile geometry it was sampled on
  DevBuf<uint8_t> mcTileEmpty;
  bool mcTileFlags = false;
  uint32_t mcTileSize[3] = {0, 0, 0}, mcTiles[3] = {0, 0, 0};
  float mcIso = 0.0f;  // isolevel of the last classify pass (its cases feed the emit)

  // profiling
  bool prof = false;
  std::string profOnly;  // when set, only launches of this name are timed
  std::vector<PendingTiming> pending;
  std::vector<hipEvent_t> eventPool;
  std::map<std::string, KernelStat> stats;
};

// Error helpers: record text, return status.
int fail(Ctx* c, int code, const std::string& what);
// Destroys the captured graph (its kernels hold addresses and constants that
// are about to change).
void dropGraph(Ctx* c);
int hipFail(Ctx* c, hipError_t e, const char* where);
#define LF_HIP(ctx, expr)                                      \
  do {                                                         \
    hipError_t _e = (expr);                                    \
    if (_e != hipSuccess) return ::lf::hipFail((ctx), _e, #expr); \
  } while (0)

// Waits for both of the context's streams (before a buffer is freed).
inline void syncStreams(Ctx* c) {
  (void)hipStreamSynchronize(c->stream);
  if (c->mcStream) (void)hipStreamSynchronize(c->mcStream);
  c->mcBusy = false;  // every event recorded so far has completed
  c->pairGuard.pending = c->pairGuardAlt.pending = false;
}

// Before a kernel that writes the current P in place on the main stream (the
// delta and finalise passes): a density field sampled from it since the grid
// build may still be reading it on the marching-cubes stream.
inline int waitPairReaders(Ctx* c) {
  if (!c->pairGuard.pending) return 0;
  c->pairGuard.pending = false;
  if (hipStreamWaitEvent(c->stream, c->pairGuard.ev, 0) != hipSuccess)
    return fail(c, LF_ERR_HIP, "waiting for the density field's read of p*");
  return 0;
}

// Health of the state (see Ctx::health): reports a raised counter of the
// pinned host copy as LF_ERR_DIVERGED; without `wait` the copy may lag the
// steps still in flight, with it the stream is synchronised first.
int checkHealth(Ctx* c, bool wait);

// After a kernel that writes P or the table on the main stream.
inline void notePositionsWritten(Ctx* c) {
  c->posEventValid = c->evPosFinal && !c->capturing &&
                     hipEventRecord(c->evPosFinal, c->stream) == hipSuccess;
}

// The main stream waits for the marching cubes on mcStream.
inline int mcJoin(Ctx* c) {
  if (!c->mcBusy) return 0;
  c->mcBusy = false;
  return hipStreamWaitEvent(c->stream, c->evMcDone, 0) == hipSuccess ? 0 : LF_ERR_HIP;
}

template <typename T> int ensure(Ctx* c, DevBuf<T>& b, size_t count, const char* what) {
  if (count <= b.cap) return LF_OK;
  if (c->capturing)  // a graph would keep the old address
    return fail(c, LF_ERR_STATE, std::string(what) + ": buffer growth during graph capture");
  syncStreams(c);
  if (b.p) {
    hipError_t e = hipFree(b.p);
    if (e != hipSuccess) return hipFail(c, e, what);
    b.p = nullptr;
    b.cap = 0;
  }
  dropGraph(c);
  size_t want = count + count / 4 + 256;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&b.p), want * sizeof(T));
  if (e != hipSuccess) return hipFail(c, e, what);
  b.cap = want;
  return LF_OK;
}
// Grows b to hold count elements keeping its first `keep` elements.
template <typename T>
int ensureKeep(Ctx* c, DevBuf<T>& b, size_t count, size_t keep, const char* what) {
  if (count <= b.cap) return LF_OK;
  if (c->capturing)
    return fail(c, LF_ERR_STATE, std::string(what) + ": buffer growth during graph capture");
  syncStreams(c);
  dropGraph(c);
  const size_t want = count + count / 4 + 256;
  T* p = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), want * sizeof(T));
  if (e != hipSuccess) return hipFail(c, e, what);
  if (b.p && keep) {
    e = hipMemcpyAsync(p, b.p, keep * sizeof(T), hipMemcpyDeviceToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      (void)hipFree(p);
      return hipFail(c, e, what);
    }
  }
  if (b.p) (void)hipFree(b.p);
  b.p = p;
  b.cap = want;
  return LF_OK;
}

template <typename T> void release(DevBuf<T>& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

// Kernel timing (profiling only): two events of the pool, which the launch
// itself stamps with the kernel's start and end (hipExtLaunchKernel: no
// marker packets of their own around the dispatch), queued for reading.
void timingEvents(Ctx* c, const char* name, PendingTiming& t);
void timingQueue(Ctx* c, PendingTiming& t);

template <typename... KArgs, typename... Args>
int launch(Ctx* c, const char* name, void (*kernel)(KArgs...), dim3 grid, dim3 block,
           size_t shmem, Args... args) {
  PendingTiming t{};
  const bool timed = c->prof && (c->profOnly.empty() || c->profOnly == name);
  const hipStream_t stream = c->launchStream ? c->launchStream : c->stream;
  if (timed) {
    timingEvents(c, name, t);
    hipExtLaunchKernelGGL(kernel, grid, block, uint32_t(shmem), stream, t.start, t.stop, 0u,
                          args...);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, shmem, stream, args...);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hipFail(c, e, name);
  if (timed) timingQueue(c, t);
  return LF_OK;
}

inline unsigned blocksFor(size_t n, unsigned per) { return unsigned((n + per - 1) / per); }

// Launch grid of a particle kernel: one workgroup per chunk, or a persistent
// grid of at most persistentBlocks workgroups when the launch takes its chunks
// from a work queue.
inline dim3 gridOf(const Ctx* c, const Work& W) {
  if (!W.queue) return dim3((W.nChunks + W.ppt - 1u) / W.ppt);
  return dim3(W.nChunks < c->persistentBlocks ? W.nChunks : c->persistentBlocks);
}

// Launches KERNEL<BS> with the context's workgroup size (128, 256 or 512
// threads; chunks hold at most that many particles).
#define LF_DISPATCH_BS(ctx, NAME, KERNEL, W, ...)                                          \
  [&]() -> int {                                                                           \
    switch ((ctx)->chunkThreads) {                                                         \
      case 128:                                                                            \
        return ::lf::launch(ctx, NAME, KERNEL<128>, ::lf::gridOf(ctx, W), dim3(128), 0, W, \
                            __VA_ARGS__);                                                  \
      case 512:                                                                            \
        return ::lf::launch(ctx, NAME, KERNEL<512>, ::lf::gridOf(ctx, W), dim3(512), 0, W, \
                            __VA_ARGS__);                                                  \
      default:                                                                             \
        return ::lf::launch(ctx, NAME, KERNEL<256>, ::lf::gridOf(ctx, W), dim3(256), 0, W, \
                            __VA_ARGS__);                                                  \
    }                                                                                      \
  }()

// ---- launchers (one .hip file each) ----------------------------------------
int prepareScenes(Ctx* c);                    // lf_api.hip: consts + chunks to the device
Work workOf(Ctx* c, int queueSlot, bool skipDone = false);  // queueSlot < 0: one workgroup per chunk
int gridBuild(Ctx* c);                        // lf_grid.hip
int pbfBuildLists(Ctx* c);                    // lf_pbf.hip
int pbfLambda(Ctx* c, bool skipDone);  // skipDone: skip the scenes that converged
int pbfDelta(Ctx* c, bool skipDone, bool final = false);  // final: finalise fused in
int pbfFinalise(Ctx* c);
int pbfXsph(Ctx* c);
int pbfDiffuse(Ctx* c);
int mcSample(Ctx* c, const lf_mc_params& m);   // lf_mc.hip
int mcClassifyCount(Ctx* c, const lf_mc_params& m);
int mcEmit(Ctx* c, const lf_mc_params& m);
int mcStaticInit(int device, uint32_t** out);  // device current
int mcMeshStatic(Ctx* c, const uint32_t S[3], const float origin[3], const float step[3],
                 float scale, float iso, const float* field, float* triPos, size_t triCap,
                 uint32_t* ntri);
int copyProbe(Ctx* c, size_t bytes, int reps, float* ms);  // lf_grid.hip
int applyDrains(Ctx* c);                      // lf_grid.hip: stable erase, updates n
int answerQueries(Ctx* c);
int setColliders(Ctx* c, size_t n, const lf_triangle* tris);
int prepareColliderGrid(Ctx* c);                    // lf_grid.hip: after the grid build
// lf_diag.hip: FP64 per-scene reductions. densityCheck runs after lambda pass
// `pass` (convergence); diagState reduces the current state into the last slot.
constexpr int kSumWidth = 8;   // count, mass, momentum x3, sum |e|, max |e|, unused
constexpr int kPassSlots = 16;
constexpr int kStatSlots = kPassSlots + 1;
int densityCheck(Ctx* c, int pass);
int diagState(Ctx* c);

}  // namespace lf// The opaque handle of the C ABI is the context itself.
struct lf_ctx : lf::Ctx {};


