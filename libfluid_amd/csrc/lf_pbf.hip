// PBF constraint passes on gfx950: neighbour lists, density/lambda, position
// correction (Jacobi), finalise, XSPH viscosity and colour diffusion.
//
// One work-item per Morton-sorted particle. The candidates of a particle are
// the particles of the 27 cells around its own cell, in the reference's order
// (oclsph_kernel.h:150-192); only those within h contribute.
//
// Neighbour lists: once per step, k_build_list records, in that visit order,
// the candidates within h + skin of each particle, measured between the
// predicted p* clamped to the bounds (P0). The lambda, delta and XSPH passes
// then walk the list with the exact r <= h test instead of the 27 cells.
//  * First lambda/delta pair (unclamped predicted p*): clamping to a box is
//    a projection onto a convex set, so |clamp(a) - clamp(b)| <= |a - b| and
//    every pair within h is listed; the lists are complete.
//  * Later passes: a candidate left out was farther than h + skin apart in
//    P0, so it can only come within h if the two particles together moved by
//    more than skin from P0. The delta pass keeps D, the largest |p* - P0| of
//    any particle; particle a walks its list only while |p*_a - P0_a| + D <
//    skin, and otherwise (or when its list overflowed) enumerates its cells.
// Either way the accepted set and its order are the reference's, so the sums
// are bit-identical to the cell walk.
//
// Buffers (sorted order): P = (p*, mass), PL = (p*, lambda), X = (world
// position, mass), V = (velocity, colour bits), T = type, P0 = clamped p* at
// the list build.
#include <algorithm>
#include <type_traits>

#include "lf_internal.hpp"

namespace lf {
namespace {

#ifndef LF_BLOCK_LAMBDA
#define LF_BLOCK_LAMBDA 8
#endif
#ifndef LF_BLOCK_DELTA
#define LF_BLOCK_DELTA 8
#endif
#ifndef LF_BLOCK_XSPH
#define LF_BLOCK_XSPH 8
#endif
// list entries in flight per lane (forEachAccepted). 8 for all three at 1M
// with the density field beside the step: 1.789 -> 1.720 ms/step over steps
// 5-25, 1.583 -> 1.488 over steps 5-105 (XSPH, 127 -> fewer VGPRs and more
// waves per SIMD, 288 -> 215 us); 16 had been faster when the passes ran
// alone (XSPH 211 -> 193 us).
constexpr int kBlockLambda = LF_BLOCK_LAMBDA;
constexpr int kBlockDelta = LF_BLOCK_DELTA;
constexpr int kBlockXsph = LF_BLOCK_XSPH;
#ifndef LF_BLOCK_BUILD
#define LF_BLOCK_BUILD 4
#endif
constexpr int kBuildBlock = LF_BLOCK_BUILD;  // candidates loaded together (list build)
#ifndef LF_REBUILD_BLOCK
#define LF_REBUILD_BLOCK 4
#endif
constexpr int kRebuildBlock = LF_REBUILD_BLOCK;  // the same, lists rebuilt by a lambda pass
// lists are padded (with the particle's own index) to whole blocks of this
// many entries, the largest block a walk reads
constexpr uint32_t kListPad = 16;
static_assert(kMaxNbr % kListPad == 0, "padding stays within the list capacity");


struct Lists {
  uint32_t* nbr;           // 24-bit entries, nbrGroupByte layout, kMaxNbr per particle
  uint32_t* count;         // entries per particle (> kMaxNbr: overflowed)
  float4* P0;              // clamped p* at the build
  uint32_t* maxDisp;       // [0]: float bits of D, the largest |p* - P0| so far;
                           // [1..3]: fallbacks of the lambda, delta, XSPH passes;
                           // [4 + j]: float bits of D_j, the largest move of any
                           // particle in delta pass j (1-based)
  uint32_t* cellDisp;      // per cell (G entries): float bits of the largest |p* - P0|
                           // of the particles of that cell
  uint8_t* epoch;          // per particle: the pass whose positions its list was
                           // built from (1: the step's build from P0)
  float4* R;               // per particle: its p* at that pass (epoch > 1)
};
constexpr int kPassSlots = 16;  // D_j kept for passes 1..kPassSlots - 1

// Whether particle a may walk its list at position (x, y, z); `fresh` marks
// the first lambda/delta pair after the build, where the lists are complete.
// A list rebuilt from the positions of pass e (lambda pass e rebuilds the
// lists of a wave that would otherwise walk its cells) holds every candidate
// within h + skin at those positions; by pass p every particle has moved at
// most sum_{j=e}^{p-1} D_j since, so the list stays complete while
// |p*_a - R_a| + that sum < skin.
__device__ inline bool listUsable(const PbfConsts& c, const Lists& L, uint32_t a, float x, float y,
                                  float z, bool fresh, int passIdx, uint32_t& cnt, uint32_t& ep) {
  cnt = L.count[a];
  ep = L.epoch[a];
  if (cnt > kMaxNbr) return false;
  if (fresh) return true;
  if (ep > 1u) {
    const float4 r = L.R[a];
    float s = __builtin_sqrtf(dist2(x - r.x, y - r.y, z - r.z));
    for (int j = int(ep); j < passIdx && j < kPassSlots; ++j) s += __uint_as_float(L.maxDisp[4 + j]);
    return passIdx < kPassSlots && s < c.skinSafeR;
  }
  const float4 p0 = L.P0[a];
  const float da = __builtin_sqrtf(dist2(x - p0.x, y - p0.y, z - p0.z));
  return da + __uint_as_float(L.maxDisp[0]) < c.skinSafe;
}

// The local form of the test: D replaced by the largest displacement among the
// particles of the 27 cells around `key` (the only possible candidates). A
// list rebuilt at pass e > 1 (reference R_a) takes twice that: a candidate b
// has moved at most |p*_b - P0_b| + |p*_b(e) - P0_b| since pass e, and the
// per-cell value is the largest |p* - P0| of the step so far.
__device__ inline bool listUsableLocal(const PbfConsts& c, const Lists& L, uint32_t a, float x,
                                       float y, float z, uint32_t key, uint32_t ep) {
  const float4 p0 = ep > 1u ? L.R[a] : L.P0[a];
  const float da = __builtin_sqrtf(dist2(x - p0.x, y - p0.y, z - p0.z));
  const uint32_t xs[3] = {dilatedDec(key, kMortonX), key & kMortonX, dilatedInc(key, kMortonX)};
  const uint32_t ys[3] = {dilatedDec(key, kMortonY), key & kMortonY, dilatedInc(key, kMortonY)};
  const uint32_t zs[3] = {dilatedDec(key, kMortonZ), key & kMortonZ, dilatedInc(key, kMortonZ)};
  // all 27 loads issued before any is used (a cell outside the table reads
  // entry 0 and is masked): one memory latency, not 27 in a row
  uint32_t d[27];
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const uint32_t code = xs[i % 3] | ys[(i / 3) % 3] | zs[i / 9];
    d[i] = L.cellDisp[c.keyBase + (code + 1u < c.G ? code : 0u)];
  }
  uint32_t m = 0;  // float bits of non-negative floats order as integers
#pragma unroll
  for (int i = 0; i < 27; ++i) {
    const uint32_t code = xs[i % 3] | ys[(i / 3) % 3] | zs[i / 9];
    m = max(m, code + 1u < c.G ? d[i] : 0u);
  }
  const float mb = __uint_as_float(m);
  return ep > 1u ? da + (mb + mb) < c.skinSafeR : da + mb < c.skinSafe;
}

// base[i] addressed as base + a 32-bit byte offset (the base stays in scalar
// registers, no 64-bit address arithmetic per gather). Particle indices are
// below 2^28 (checked at upload), so i * 16 fits.
template <typename T> __device__ inline T at(const T* __restrict__ base, uint32_t i) {
  static_assert(sizeof(T) <= 16, "records of at most 16 bytes");
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * uint32_t(sizeof(T)));
}

// Gathered neighbour records: a float4 whose xyz is p*, position and velocity
// together (XSPH), or p* and its fixed-point record (LF_FLAG_PACKED16 lists
// rebuilt by the lambda pass).
struct PosVel {
  float4 p, v;
};
struct PosQ {
  float4 p;
  uint2 q;
};
__device__ inline const float4& posOf(const float4& r) { return r; }
__device__ inline const float4& posOf(const PosVel& r) { return r.p; }
__device__ inline const float4& posOf(const PosQ& r) { return r.p; }

// Builds particle a's list in the reference's visit order (the 27 cells of
// key, in sorted order within a cell): the candidates whose position posOf(
// load(b)) lies within h + skin of pa. visit(b, record, valid) runs for each
// candidate in that order (valid false for the padding of a block), so a pass
// that rebuilds lists evaluates its pair terms in the same walk. Each lane
// collects its accepted candidates in a 16-entry LDS ring (win, two 8-entry
// groups) and, once per block of candidates, stores a completed group as 24
// bytes of its list: a block adds at most BB <= 8 entries, so at most
// one group completes in it and the group still to be stored is never
// overwritten. (Checking after every candidate ran the store path on almost
// every step: with 64 lanes, some lane nearly always completes a group.) The
// list is padded with a itself to whole blocks of kListPad entries; the
// count (kMaxNbr + 1: overflowed, or a candidate 2^23 positions away) goes
// to L.count[a].
template <int BS, int BB = kBuildBlock, typename Load, typename Visit>
__device__ inline void buildList(const PbfConsts& c, float listR2, const Lists& L,
                                 uint32_t* __restrict__ win, uint32_t a, float4 pa, uint32_t key,
                                 const uint32_t* __restrict__ table, int checkFar, Load&& load,
                                 Visit&& visit) {
  constexpr uint32_t kRing = 2 * kNbrGroup;
  // at most 7 entries wait before a block, so at most 15 after it: one group
  // completes at most, and a block's writes always find a free slot
  static_assert(BB <= int(kNbrGroup), "a block completes at most one group");
  uint32_t cnt = 0, stored = 0;  // entries accepted / written to the list (a multiple of 8)
  bool far = false;  // a candidate 2^23 or more sorted positions away: no list
  // group i / 8 of a's list starts at nbrGroupByte(a, 0) + (i / 8) * 1536 bytes
  char* const list0 = reinterpret_cast<char*>(L.nbr) + nbrGroupByte(a, 0);
  static_assert(nbrGroupByte(0, 1) == kNbrWaveGroupBytes, "groups of a particle 1536 bytes apart");
  const uint32_t bias = kNbrBias - a;  // ring entries are stored biased: b - a + 2^23
  auto storeGroup = [&](uint32_t first, uint32_t valid) {  // entries [first, first + 8)
    // first is a multiple of 8: its slots are (first & 8) + 0..7, no wrap
    const uint32_t* w = win + (first & kNbrGroup) * BS + threadIdx.x;
    uint32_t e[kNbrGroup], q[6];
#pragma unroll
    for (uint32_t j = 0; j < kNbrGroup; ++j) e[j] = j < valid ? w[j * BS] : kNbrBias;
    packGroup(e, q);
    uint3* g = reinterpret_cast<uint3*>(list0 + (first / kNbrGroup) * kNbrWaveGroupBytes);
    g[0] = make_uint3(q[0], q[1], q[2]);
    g[1] = make_uint3(q[3], q[4], q[5]);
  };
  forEachNeighbourCellRolled(key, table, c.G, [&](uint32_t s, uint32_t e) {
    // a cell's candidates are the run [s, e): its ends bound their distance
    if (checkFar && ((s - a + kNbrBias) >> 24 || (e - 1u - a + kNbrBias) >> 24)) far = true;
    for (uint32_t b0 = s; b0 < e; b0 += BB) {
      decltype(load(0u)) rec[BB];  // candidates loaded together (indices clamped into the cell)
#pragma unroll
      for (int k = 0; k < BB; ++k) rec[k] = load(min(b0 + k, e - 1u));
#pragma unroll
      for (int k = 0; k < BB; ++k) {
        const uint32_t b = b0 + k;
        const float4& pb = posOf(rec[k]);
        // both tests always evaluated (the load index is clamped into the
        // cell): no branch per candidate
        const uint32_t accept =
            (b < e ? 1u : 0u) &
            (dist2(pa.x - pb.x, pa.y - pb.y, pa.z - pb.z) <= listR2 ? 1u : 0u);
        // written whether accepted or not: slot cnt % 16 is the next free
        // one, so a rejected candidate is overwritten by the next accepted
        win[(cnt % kRing) * BS + threadIdx.x] = b + bias;
        cnt += accept;
        visit(b, rec[k], b < e);
      }
      if (cnt - stored >= kNbrGroup) {
        if (stored + kNbrGroup <= kMaxNbr) storeGroup(stored, kNbrGroup);
        stored += kNbrGroup;
      }
    }
  });
  // The last cnt % 8 entries, and the list padded with a itself to whole
  // blocks of kListPad entries: readers load whole blocks and skip the
  // entries past the count, so every loaded index is a valid particle.
  if (cnt <= kMaxNbr && !far) {
    if (cnt != stored) storeGroup(stored, cnt - stored);
    for (uint32_t s = (cnt + kNbrGroup - 1) / kNbrGroup * kNbrGroup; s % kListPad != 0u;
         s += kNbrGroup)
      storeGroup(s, 0);
  }
  L.count[a] = far ? kMaxNbr + 1u : cnt;
}

// Neighbour lists of the predicted p* (see the file comment), from P0 (the
// clamped p*, written by the reorder kernel): epoch 1 for every particle.
template <int BS>
__global__ __launch_bounds__(BS) void k_build_list(Work W, const uint32_t* __restrict__ keys,
                                                   const uint32_t* __restrict__ table,
                                                   const uint8_t* __restrict__ T, Lists L,
                                                   int checkFar) {
  // win[j * BS + lane] holds entry j mod 16 of the lane's list; lanes side by
  // side, so no bank conflicts
  __shared__ uint32_t win[2 * kNbrGroup * BS];
  if (blockIdx.x == 0 && threadIdx.x < 4 + kPassSlots) L.maxDisp[threadIdx.x] = 0u;
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    const uint32_t a = ch.begin + threadIdx.x;
    if (a >= ch.end) return;
    L.epoch[a] = 1u;
    if (T[a] == LF_OBSTACLE) {
      L.count[a] = 0;
      return;
    }
    buildList<BS>(c, c.listR2, L, win, a, L.P0[a], keys[a], table + c.keyBase, checkFar,
                  [&](uint32_t b) { return at(L.P0, b); }, [](uint32_t, const float4&, bool) {});
  });
}

// Accumulators of the lambda pass; KAHAN adds Neumaier compensation terms.
struct LambdaSums {
  float gx = 0.0f, gy = 0.0f, gz = 0.0f, rho = 0.0f;
  float cgx = 0.0f, cgy = 0.0f, cgz = 0.0f, crho = 0.0f;
};

__device__ inline void neumaier(float& sum, float& comp, float term) {
  const float t = sum + term;
  comp += (fabsf(sum) >= fabsf(term)) ? (sum - t) + term : (term - t) + sum;
  sum = t;
}

// One accepted pair of sph_lambda (oclsph_kernel.h:324-326): gradient sum
// g += gradW / rho0 and density rho += m_a W6(r).
template <bool KAHAN, int AR>
__device__ inline void lambdaTerm(const PbfConsts& c, float ma, float dx, float dy, float dz,
                                  float d2, LambdaSums& A) {
  const float r = sqrtM<AR>(d2);
  const float w = poly6W(c, r);
  const float sc = spikyScaleM<AR>(c, r);
  if (!KAHAN) {
    A.gx = __builtin_fmaf(dx * sc, c.rhoRecip, A.gx);
    A.gy = __builtin_fmaf(dy * sc, c.rhoRecip, A.gy);
    A.gz = __builtin_fmaf(dz * sc, c.rhoRecip, A.gz);
    A.rho = __builtin_fmaf(ma, w, A.rho);
  } else {
    neumaier(A.gx, A.cgx, (dx * sc) * c.rhoRecip);
    neumaier(A.gy, A.cgy, (dy * sc) * c.rhoRecip);
    neumaier(A.gz, A.cgz, (dz * sc) * c.rhoRecip);
    neumaier(A.rho, A.crho, ma * w);
  }
}


// Block i of particle a's list (entries i .. i + BLK - 1, groups of 8) in two
// halves: the packed words (two 12-byte loads per group), and their entries
// as particle indices, b = a + entry - 2^23. A walk loads the
// next block's words while the current block's gathers are in flight, so a
// block costs one memory latency instead of two (the list load, then the
// gathers it addresses).
constexpr int kListWords = 6;  // words per group of 8 entries
template <int BLK>
__device__ inline void loadListWords(const Lists& L, uint32_t a, uint32_t i,
                                     uint32_t w[BLK / kNbrGroup * kListWords]) {
  const char* const list0 = reinterpret_cast<const char*>(L.nbr) + nbrGroupByte(a, 0);
#pragma unroll
  for (int g = 0; g < BLK / int(kNbrGroup); ++g) {
    const uint3* q =
        reinterpret_cast<const uint3*>(list0 + (i / kNbrGroup + g) * kNbrWaveGroupBytes);
    const uint3 lo = q[0], hi = q[1];
    w[g * kListWords + 0] = lo.x;
    w[g * kListWords + 1] = lo.y;
    w[g * kListWords + 2] = lo.z;
    w[g * kListWords + 3] = hi.x;
    w[g * kListWords + 4] = hi.y;
    w[g * kListWords + 5] = hi.z;
  }
}
template <int BLK>
__device__ inline void decodeListWords(uint32_t a, const uint32_t w[BLK / kNbrGroup * kListWords],
                                       uint32_t b[BLK]) {
  const uint32_t base = a - kNbrBias;
#pragma unroll
  for (int g = 0; g < BLK / int(kNbrGroup); ++g) {
    uint32_t e[kNbrGroup];
    unpackGroup(w + g * kListWords, e);
#pragma unroll
    for (int j = 0; j < int(kNbrGroup); ++j) b[g * kNbrGroup + j] = base + e[j];
  }
}

// The 27-cell walk of a wave whose lists are not usable: candidates loaded
// kWalkBlock at a time (indices clamped into the cell, so every load is a
// valid particle) and used in order, with the next cell's table entries
// loaded during the visit; one memory latency per block, not per candidate.
// rel(record, dx, dy, dz) gives p_a - p_b from a gathered record: FP32
// differences of the float4 p* (RelF32), or of the fixed-point coordinates
// (RelQ, LF_FLAG_PACKED16); use(b, record, dx, dy, dz, d2) runs for the
// accepted ones.
struct RelF32 {
  float4 pa;
  template <typename R> __device__ void operator()(const R& r, float& dx, float& dy, float& dz) const {
    const float4& q = posOf(r);
    dx = pa.x - q.x;
    dy = pa.y - q.y;
    dz = pa.z - q.z;
  }
};
struct QRec {  // a neighbour of the packed delta pass: its record and its lambda
  uint2 q;
  float lam;
};
__device__ inline uint2 qOf(const uint2& r) { return r; }
__device__ inline uint2 qOf(const QRec& r) { return r.q; }
__device__ inline uint2 qOf(const PosQ& r) { return r.q; }
struct RelQ {
  QPos qa;
  float unit;
  template <typename R> __device__ void operator()(const R& r, float& dx, float& dy, float& dz) const {
    const QPos qb = unpackPos(qOf(r));
    dx = qDiff(qa.x, qb.x, unit);
    dy = qDiff(qa.y, qb.y, unit);
    dz = qDiff(qa.z, qb.z, unit);
  }
};

constexpr int kWalkBlock = 4;
template <typename Load, typename Rel, typename Use>
__device__ inline void walkCells(const PbfConsts& c, uint32_t key,
                                 const uint32_t* __restrict__ table, Load&& load, Rel&& rel,
                                 Use&& use) {
  forEachNeighbourCellRolled(key, table, c.G, [&](uint32_t s, uint32_t e) {
    for (uint32_t b0 = s; b0 < e; b0 += kWalkBlock) {
      decltype(load(0u)) p[kWalkBlock];
#pragma unroll
      for (int k = 0; k < kWalkBlock; ++k) p[k] = load(min(b0 + k, e - 1u));
#pragma unroll
      for (int k = 0; k < kWalkBlock; ++k) {
        float dx, dy, dz;
        rel(p[k], dx, dy, dz);
        const float d2 = dist2(dx, dy, dz);
        if (b0 + k < e && d2 <= c.rmax2) use(b0 + k, p[k], dx, dy, dz, d2);
      }
    }
  });
}

#ifndef LF_PAIR_TILES
#define LF_PAIR_TILES 0
#endif
// ---- list-free pair passes over LDS cell tiles (a measured alternative) ----
// Built with LF_PAIR_TILES=1: the lambda and delta passes (plain FP32) stage
// the particles of their chunk's halo cells (the chunk's cell box +-1 per
// axis) in LDS, x fastest, with a local cell table, and each lane walks its
// 27 cells there: the same candidates in the same order as walkCells, so the
// sums are bit-identical. A chunk whose halo does not fit, or touches the
// table's edge, walks its cells in global memory.
constexpr uint32_t kTileHaloCells = 1024;
constexpr uint32_t kTileCap = 2048;  // staged records (32 KB)

struct HaloTile {
  uint32_t lo[3];  // cell coordinates of the halo's first cell
  uint32_t hx, hy;
  bool ok;
};

[[maybe_unused]] __device__ inline uint32_t compact10(uint32_t v) {  // inverse of spread10
  v &= 0x09249249u;
  v = (v | (v >> 2)) & 0x030C30C3u;
  v = (v | (v >> 4)) & 0x0300F00Fu;
  v = (v | (v >> 8)) & 0x030000FFu;
  v = (v | (v >> 16)) & 0x000003FFu;
  return v;
}

// Called by every thread of the workgroup for chunk ch (scene-local keys and
// table); src: the records the pass gathers (16 bytes each).
template <int BS>
__device__ inline HaloTile stageHalo(const PbfConsts& c, const Chunk& ch,
                                     const uint32_t* __restrict__ keys,
                                     const uint32_t* __restrict__ table,
                                     const float4* __restrict__ src, float4* rec, uint32_t* start,
                                     uint32_t* red) {
  constexpr uint32_t kW = BS / 64;
  const uint32_t a = ch.begin + threadIdx.x;
  const bool valid = a < ch.end;
  const uint32_t key = valid ? keys[a] : 0u;
  const uint32_t x = compact10(key), y = compact10(key >> 1), z = compact10(key >> 2);
  uint32_t mn[3] = {valid ? x : 1023u, valid ? y : 1023u, valid ? z : 1023u};
  uint32_t mx[3] = {valid ? x : 0u, valid ? y : 0u, valid ? z : 0u};
  uint32_t bad = (valid && !(key < c.G - 1u && c.G > 0u)) ? 1u : 0u;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      mn[k] = min(mn[k], uint32_t(__shfl_xor(int(mn[k]), o, 64)));
      mx[k] = max(mx[k], uint32_t(__shfl_xor(int(mx[k]), o, 64)));
    }
    bad |= uint32_t(__shfl_xor(int(bad), o, 64));
  }
  const uint32_t w = threadIdx.x >> 6;
  if (__lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      red[w * 8 + k] = mn[k];
      red[w * 8 + 3 + k] = mx[k];
    }
    red[w * 8 + 6] = bad;
  }
  __syncthreads();
  HaloTile t{};
  uint32_t hi[3];
  bad = 0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    t.lo[k] = 1023u;
    hi[k] = 0u;
  }
  for (uint32_t i = 0; i < kW; ++i) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      t.lo[k] = min(t.lo[k], red[i * 8 + k]);
      hi[k] = max(hi[k], red[i * 8 + 3 + k]);
    }
    bad |= red[i * 8 + 6];
  }
  bool ok = bad == 0u;
#pragma unroll
  for (int k = 0; k < 3; ++k) ok = ok && t.lo[k] >= 1u && hi[k] <= 1022u;
  t.hx = ok ? hi[0] - t.lo[0] + 3u : 0u;
  t.hy = ok ? hi[1] - t.lo[1] + 3u : 0u;
  const uint32_t hz = ok ? hi[2] - t.lo[2] + 3u : 0u;
  const uint32_t nh = t.hx * t.hy * hz;
  ok = ok && nh <= kTileHaloCells;
#pragma unroll
  for (int k = 0; k < 3; ++k) t.lo[k] -= 1u;  // the halo's first cell
  // per-thread runs of halo cells: counts, then an exclusive scan
  const uint32_t per = ok ? (nh + BS - 1u) / BS : 0u;
  const uint32_t h0 = min(threadIdx.x * per, nh), h1 = min(h0 + per, nh);
  auto codeOf = [&](uint32_t h) {
    return mortonEncode(t.lo[0] + h % t.hx, t.lo[1] + (h / t.hx) % t.hy, t.lo[2] + h / (t.hx * t.hy));
  };
  uint32_t sum = 0;
  for (uint32_t h = h0; h < h1; ++h) {
    const uint32_t code = codeOf(h);
    sum += code + 1u < c.G ? table[code + 1u] - table[code] : 0u;
  }
  uint32_t xs = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t yv = uint32_t(__shfl_up(int(xs), o, 64));
    if (__lane_id() >= uint32_t(o)) xs += yv;
  }
  __syncthreads();  // red is reused
  if (__lane_id() == 63) red[w] = xs;
  __syncthreads();
  uint32_t run = xs - sum, total = 0;
  for (uint32_t i = 0; i < kW; ++i) {
    if (i < w) run += red[i];
    total += red[i];
  }
  ok = ok && total <= kTileCap;
  if (ok) {
    for (uint32_t h = h0; h < h1; ++h) {
      const uint32_t code = codeOf(h);
      start[h] = run;
      if (code + 1u < c.G) {
        const uint32_t s0 = table[code], s1 = table[code + 1u];
        for (uint32_t b = s0; b < s1; ++b) rec[run + (b - s0)] = src[b];
        run += s1 - s0;
      }
    }
    if (threadIdx.x == BS - 1) start[nh] = total;
  }
  __syncthreads();
  t.ok = ok;
  return t;
}

// Particle a's accepted neighbours from the staged halo (t.ok), in the
// reference order: z rows outer, then y rows, each row the x-run of three
// consecutive halo cells.
template <typename Rel, typename Use>
__device__ inline void walkHalo(const PbfConsts& c, const HaloTile& t, uint32_t key,
                                const float4* rec, const uint32_t* start, Rel&& rel, Use&& use) {
  const uint32_t bx = compact10(key) - t.lo[0] - 1u, by = compact10(key >> 1) - t.lo[1] - 1u,
                 bz = compact10(key >> 2) - t.lo[2] - 1u;  // halo index of the -1 neighbours
  for (uint32_t dz = 0; dz < 3; ++dz)
    for (uint32_t dy = 0; dy < 3; ++dy) {
      const uint32_t row = ((bz + dz) * t.hy + (by + dy)) * t.hx + bx;
      const uint32_t e = start[row + 3];
      for (uint32_t b = start[row]; b < e; ++b) {
        const float4 pb = rec[b];
        float dx, dy2, dz2;
        rel(pb, dx, dy2, dz2);
        const float d2 = dist2(dx, dy2, dz2);
        if (d2 <= c.rmax2) use(b, pb, dx, dy2, dz2, d2);
      }
    }
}

// Visits the accepted neighbours (r <= h) of particle a at position pa, in
// the reference's order: through its list when usable, else the 27 cells.
// load(b) returns the candidate's float4 (xyz = p*). The list is walked in
// blocks of BLK entries: BLK index loads, then BLK gathers, then the BLK pair
// terms in list order, so a block costs one memory latency instead of BLK.
// Entries past the count (the tail of the last block) hold a itself (the
// build pads every list to whole blocks), a valid address, and are skipped.
//
// Rebuild (the lambda passes after the first, win non-null): a wave that
// would walk its cells rebuilds its lanes' lists from this pass's positions
// in the same walk (buildList over the records of loadPos, which gives p* and
// the pass's record), evaluating the pair terms as it goes, and marks them
// epoch passIdx. The delta pass of the same iteration reads the same
// positions, so it walks the new lists, and later passes keep them while
// listUsable allows. The accepted neighbours and their order are the cell
// walk's.
template <int BS, int BLK, bool REBUILD, typename Load, typename Rel, typename Use,
          typename LoadPos>
__device__ inline void forEachAccepted(const PbfConsts& c, const Lists& L, int pass, bool fresh,
                                       int passIdx, uint32_t a, float4 pa, uint32_t key,
                                       const uint32_t* __restrict__ table, Load&& load, Rel&& rel,
                                       Use&& use, uint32_t* __restrict__ win, int checkFar,
                                       LoadPos&& loadPos) {
  uint32_t cnt, ep;
  bool usable = listUsable(c, L, a, pa.x, pa.y, pa.z, fresh, passIdx, cnt, ep);
  if (!usable && cnt <= kMaxNbr && (ep == 1u || passIdx < kPassSlots))
    usable = listUsableLocal(c, L, a, pa.x, pa.y, pa.z, key, ep);
  const uint64_t fell = __ballot(!usable);
  if (fell && __lane_id() == __builtin_ctzll(fell)) atomicAdd(&L.maxDisp[pass], uint32_t(__popcll(fell)));
  // A wave with any lane walking its cells walks cells on every lane: the
  // walk takes as long for one lane as for 64, and list + walk in sequence
  // cost more than the walk alone. The accepted neighbours and their order
  // are the same either way.
  if (fell) usable = false;
  if constexpr (REBUILD) {
    if (!usable && win && passIdx < kPassSlots) {
      buildList<BS, kRebuildBlock>(c, c.listR2R, L, win, a, pa, key, table, checkFar, loadPos,
                    [&](uint32_t b, const auto& r, bool valid) {
                      float dx, dy, dz;
                      rel(r, dx, dy, dz);
                      const float d2 = dist2(dx, dy, dz);
                      if (valid && d2 <= c.rmax2) use(b, r, dx, dy, dz, d2);
                    });
      L.epoch[a] = uint8_t(passIdx);
      L.R[a] = make_float4(pa.x, pa.y, pa.z, 0.0f);
      return;
    }
  }
  if (usable) {
    static_assert(BLK % kNbrGroup == 0, "blocks of whole list groups");
    static_assert(kListPad % BLK == 0, "the build pads lists to whole blocks");
    constexpr int kW = BLK / int(kNbrGroup) * kListWords;
    uint32_t w[kW];
    if (cnt > 0) loadListWords<BLK>(L, a, 0, w);
    for (uint32_t i = 0; i < cnt; i += BLK) {
      uint32_t b[BLK];
      decodeListWords<BLK>(a, w, b);  // entries past cnt hold a
      decltype(load(0u)) p[BLK];
#pragma unroll
      for (int k = 0; k < BLK; ++k) p[k] = load(b[k]);
      if (i + BLK < cnt) loadListWords<BLK>(L, a, i + BLK, w);  // the next block's entries
#pragma unroll
      for (int k = 0; k < BLK; ++k) {
        float dx, dy, dz;
        rel(p[k], dx, dy, dz);
        const float d2 = dist2(dx, dy, dz);
        if (i + k < cnt && d2 <= c.rmax2) use(b[k], p[k], dx, dy, dz, d2);
      }
    }
  } else {
    walkCells(c, key, table, load, rel, use);
  }
}

#ifndef LF_XSPH_SPLIT
#define LF_XSPH_SPLIT 1
#endif
// forEachAccepted for a pass that needs a second record (V) of its accepted
// neighbours only: each block gathers the positions, then V of the entries
// within h (lanes whose entry lies in the skin issue no request), then runs
// use(p_b, v_b, d2) in list order. Same neighbours, same order.
template <int BLK, typename Use>
__device__ inline void forEachAcceptedSplit(const PbfConsts& c, const Lists& L, int pass,
                                            int passIdx, uint32_t a, float4 pa, uint32_t key,
                                            const uint32_t* __restrict__ table,
                                            const float4* __restrict__ P,
                                            const float4* __restrict__ V, Use&& use) {
  uint32_t cnt, ep;
  bool usable = listUsable(c, L, a, pa.x, pa.y, pa.z, false, passIdx, cnt, ep);
  if (!usable && cnt <= kMaxNbr && (ep == 1u || passIdx < kPassSlots))
    usable = listUsableLocal(c, L, a, pa.x, pa.y, pa.z, key, ep);
  const uint64_t fell = __ballot(!usable);
  if (fell && __lane_id() == __builtin_ctzll(fell)) atomicAdd(&L.maxDisp[pass], uint32_t(__popcll(fell)));
  if (fell) usable = false;  // the whole wave walks cells (as forEachAccepted)
  if (usable) {
    static_assert(BLK % kNbrGroup == 0, "blocks of whole list groups");
    static_assert(kListPad % BLK == 0, "the build pads lists to whole blocks");
    constexpr int kW = BLK / int(kNbrGroup) * kListWords;
    uint32_t w[kW];
    if (cnt > 0) loadListWords<BLK>(L, a, 0, w);
    for (uint32_t i = 0; i < cnt; i += BLK) {
      uint32_t b[BLK];
      decodeListWords<BLK>(a, w, b);  // entries past cnt hold a
      float4 p[BLK];
#pragma unroll
      for (int k = 0; k < BLK; ++k) p[k] = at(P, b[k]);
      if (i + BLK < cnt) loadListWords<BLK>(L, a, i + BLK, w);  // the next block's entries
      float d2[BLK];
      bool in[BLK];
      float4 v[BLK];
#pragma unroll
      for (int k = 0; k < BLK; ++k) {
        d2[k] = dist2(pa.x - p[k].x, pa.y - p[k].y, pa.z - p[k].z);
        in[k] = i + k < cnt && d2[k] <= c.rmax2;
        v[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (in[k]) v[k] = at(V, b[k]);
      }
#pragma unroll
      for (int k = 0; k < BLK; ++k)
        if (in[k]) use(p[k], v[k], d2[k]);
    }
  } else {
    walkCells(
        c, key, table, [&](uint32_t b) { return PosVel{at(P, b), at(V, b)}; }, RelF32{pa},
        [&](uint32_t, const PosVel& r, float, float, float, float d2) { use(r.p, r.v, d2); });
  }
}

// sph_lambda (oclsph_kernel.h:305-332). Rejected candidates contribute
// exactly zero in the reference, so walking only the accepted ones gives the
// same sums in the same order.
// PACKED (LF_FLAG_PACKED16): neighbours' positions from the fixed-point
// records Q; the pass copies each record to QL and lambda to LAM for the
// delta pass (as PL for the FP32 gathers).
template <int BS, bool KAHAN, int AR, bool PACKED, bool TILES = false>
__global__ __launch_bounds__(BS) void k_lambda(Work W, const uint32_t* __restrict__ keys,
                                               const uint32_t* __restrict__ table,
                                               const uint8_t* __restrict__ T,
                                               const float4* __restrict__ P,
                                               float4* __restrict__ PL, float* __restrict__ rhoOut,
                                               Lists L, int fresh, int passIdx, int rebuild,
                                               int checkFar, const uint2* __restrict__ Q,
                                               uint2* __restrict__ QL, float* __restrict__ LAM) {
  // the ring of list rebuilds (forEachAccepted), 2 * kNbrGroup * BS words:
  // dynamic LDS, given only to the passes that may rebuild, so that the
  // first pass leaves that LDS to the kernels beside it
  extern __shared__ uint32_t win[];
  __shared__ float4 haloRec[TILES ? kTileCap : 1];
  __shared__ uint32_t haloStart[TILES ? kTileHaloCells + 1 : 1];
  __shared__ uint32_t haloRed[TILES ? 8 * (BS / 64) : 1];
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    HaloTile tile{};
    if constexpr (TILES)
      tile = stageHalo<BS>(c, ch, keys, table + c.keyBase, P, haloRec, haloStart, haloRed);
    const uint32_t a = ch.begin + threadIdx.x;
    if (a >= ch.end) return;
    const float4 pa = P[a];
    uint2 qa = make_uint2(0u, 0u);
    if (PACKED) {
      qa = Q[a];
      QL[a] = qa;
    }
    if (T[a] == LF_OBSTACLE) {
      PL[a] = make_float4(pa.x, pa.y, pa.z, 0.0f);
      rhoOut[a] = 0.0f;
      if (PACKED) LAM[a] = 0.0f;
      return;
    }
    LambdaSums A;
    auto term = [&](uint32_t, const auto&, float dx, float dy, float dz, float d2) {
      lambdaTerm<KAHAN, AR>(c, pa.w, dx, dy, dz, d2, A);
    };
    uint32_t* const ring = (rebuild && !TILES) ? win : nullptr;
    if constexpr (TILES) {
      if (tile.ok)
        walkHalo(c, tile, keys[a], haloRec, haloStart, RelF32{pa}, term);
      else
        walkCells(c, keys[a], table + c.keyBase, [&](uint32_t b) { return at(P, b); }, RelF32{pa},
                  term);
    } else if (PACKED)
      forEachAccepted<BS, kBlockLambda, true>(
          c, L, 1, fresh != 0, passIdx, a, pa, keys[a], table + c.keyBase,
          [&](uint32_t b) { return at(Q, b); }, RelQ{unpackPos(qa), c.qUnit}, term, ring,
          checkFar, [&](uint32_t b) { return PosQ{at(P, b), at(Q, b)}; });
    else
      forEachAccepted<BS, kBlockLambda, true>(
          c, L, 1, fresh != 0, passIdx, a, pa, keys[a], table + c.keyBase,
          [&](uint32_t b) { return at(P, b); }, RelF32{pa}, term, ring, checkFar,
          [&](uint32_t b) { return at(P, b); });
    if (KAHAN) {
      A.gx += A.cgx;
      A.gy += A.cgy;
      A.gz += A.cgz;
      A.rho += A.crho;
    }
    const float norm2 = dist2(A.gx, A.gy, A.gz);
    const float C1 = A.rho / c.rho0 - 1.0f;
    const float lam = -C1 / (norm2 + c.cfm);
    PL[a] = make_float4(pa.x, pa.y, pa.z, lam);
    rhoOut[a] = A.rho;
    if (PACKED) LAM[a] = lam;
  });
}

// One accepted pair of sph_delta (oclsph_kernel.h:355-358).
template <int AR>
__device__ inline void deltaTerm(const PbfConsts& c, float la, float lb, float dx, float dy,
                                 float dz, float d2, float& ax, float& ay, float& az) {
  const float r = sqrtM<AR>(d2);
  const float q = poly6W(c, r) * c.wdqRecip;
  const float corr = -c.corrK * ((q * q) * (q * q));
  const float factor = ((la + lb) + corr) * c.rhoRecip;
  const float sc = spikyScaleM<AR>(c, r);
  ax = __builtin_fmaf(dx * sc, factor, ax);
  ay = __builtin_fmaf(dy * sc, factor, ay);
  az = __builtin_fmaf(dz * sc, factor, az);
}

// sph_delta (oclsph_kernel.h:335-384) with the mesh colliders; Jacobi: reads PL,
// writes Pout. deltaP is not stored (nothing reads it after the pass). Also
// raises D and the per-cell displacement to the new positions' (one atomic
// per wave, and per run of equal keys).
//
// FINAL (the last pass of a step whose passes all run): also does
// k_finalise's work for the particle, from the same values in the same
// order, so the separate finalise launch and its pass over the state go.
template <int BS, int AR, bool FINAL, bool PACKED, bool TILES = false>
__global__ __launch_bounds__(BS) void k_delta(Work W, const uint32_t* __restrict__ keys,
                                              const uint32_t* __restrict__ table,
                                              const uint8_t* __restrict__ T,
                                              const float4* __restrict__ PL,
                                              float4* __restrict__ X, Lists L, int fresh,
                                              float4* __restrict__ Pout, float4* __restrict__ V,
                                              ColliderGrid mesh,
                                              const uint32_t* __restrict__ diffused,
                                              int useDiffused, const uint2* __restrict__ QL,
                                              const float* __restrict__ LAM,
                                              uint2* __restrict__ Qout, int passIdx) {
  __shared__ float waveMax[BS / 64], waveMove[BS / 64];
  __shared__ float4 haloRec[TILES ? kTileCap : 1];
  __shared__ uint32_t haloStart[TILES ? kTileHaloCells + 1 : 1];
  __shared__ uint32_t haloRed[TILES ? 8 * (BS / 64) : 1];
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    HaloTile tile{};
    if constexpr (TILES)
      tile = stageHalo<BS>(c, ch, keys, table + c.keyBase, PL, haloRec, haloStart, haloRed);
    const uint32_t a = ch.begin + threadIdx.x;
    const bool valid = a < ch.end;
    float disp = 0.0f;  // |p* - P0|, since the step's list build
    float move = 0.0f;  // |p* - p* of this pass's input|: this pass's move
    if (valid) {
      const float4 pa = PL[a];
      const float mass = X[a].w;
      if (T[a] == LF_OBSTACLE) {
        // finalise leaves obstacles in place with P.w = 0 (no fluid mass)
        Pout[a] = make_float4(pa.x, pa.y, pa.z, FINAL ? 0.0f : mass);
        if (PACKED) Qout[a] = QL[a];
      } else {
        const float la = pa.w;
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        auto noRebuild = [](uint32_t) { return make_float4(0.0f, 0.0f, 0.0f, 0.0f); };
        auto dterm = [&](uint32_t, const float4& pb, float dx, float dy, float dz, float d2) {
          deltaTerm<AR>(c, la, pb.w, dx, dy, dz, d2, ax, ay, az);
        };
        if constexpr (TILES) {
          if (tile.ok)
            walkHalo(c, tile, keys[a], haloRec, haloStart, RelF32{pa}, dterm);
          else
            walkCells(c, keys[a], table + c.keyBase, [&](uint32_t b) { return at(PL, b); },
                      RelF32{pa}, dterm);
        } else if (PACKED)
          forEachAccepted<BS, kBlockDelta, false>(
              c, L, 2, fresh != 0, passIdx, a, pa, keys[a], table + c.keyBase,
              [&](uint32_t b) { return QRec{at(QL, b), at(LAM, b)}; },
              RelQ{unpackPos(QL[a]), c.qUnit},
              [&](uint32_t, const QRec& r, float dx, float dy, float dz, float d2) {
                deltaTerm<AR>(c, la, r.lam, dx, dy, dz, d2, ax, ay, az);
              },
              nullptr, 0, noRebuild);
        else
          forEachAccepted<BS, kBlockDelta, false>(
              c, L, 2, fresh != 0, passIdx, a, pa, keys[a], table + c.keyBase,
              [&](uint32_t b) { return at(PL, b); }, RelF32{pa},
              [&](uint32_t, const float4& pb, float dx, float dy, float dz, float d2) {
                deltaTerm<AR>(c, la, pb.w, dx, dy, dz, d2, ax, ay, az);
              },
              nullptr, 0, noRebuild);
        // pos = (p* + dp) * scale, mesh colliders (:364-368), clamped to the
        // bounds (:371), p* = pos/scale (:374)
        const float acc[3] = {ax, ay, az};
        const float p[3] = {pa.x, pa.y, pa.z};
        float pos[3], o[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) pos[k] = (p[k] + acc[k]) * c.scale;
        if (mesh.n) {
          const float4 va = V[a];
          float vel[3] = {va.x, va.y, va.z};
          if (collide(mesh, pos, vel))
            V[a] = make_float4(vel[0], vel[1], vel[2], va.w);  // velocity[a] = resp.velocity
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          float q = pos[k];
          q = (c.minB[k] < q) ? q : c.minB[k];
          q = (q < c.maxB[k]) ? q : c.maxB[k];
          o[k] = q / c.scale;
        }
        Pout[a] = make_float4(o[0], o[1], o[2], mass);
        if (PACKED) Qout[a] = packPos(c, o[0], o[1], o[2]);
        if (FINAL) {  // k_finalise on p* = o; V[a] after the collider response
          const float4 x = X[a];
          const float4 v = V[a];
          const float dX = o[0] - x.x / c.scale;
          const float dY = o[1] - x.y / c.scale;
          const float dZ = o[2] - x.z / c.scale;
          X[a] = make_float4(o[0] * c.scale, o[1] * c.scale, o[2] * c.scale, x.w);
          const float col = useDiffused ? __uint_as_float(diffused[a]) : v.w;
          V[a] = make_float4(__builtin_fmaf(dX, c.invDt, v.x) * c.vd,
                             __builtin_fmaf(dY, c.invDt, v.y) * c.vd,
                             __builtin_fmaf(dZ, c.invDt, v.z) * c.vd, col);
        }
        const float4 p0 = L.P0[a];
        disp = __builtin_sqrtf(dist2(o[0] - p0.x, o[1] - p0.y, o[2] - p0.z));
        if (!(disp < 3.0e38f)) disp = 3.0e38f;  // NaN / inf: every list unusable
        move = __builtin_sqrtf(dist2(o[0] - pa.x, o[1] - pa.y, o[2] - pa.z));
        if (!(move < 3.0e38f)) move = 3.0e38f;
      }
    }
    // per-cell maximum: one atomic per run of equal keys within the wave
    const uint32_t key = valid ? keys[a] : 0xFFFFFFFFu;
    const uint32_t prevKey = __shfl_up(key, 1, 64);
    float runMax = disp;
    const bool head = __lane_id() == 0 || prevKey != key;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const float other = __shfl_down(runMax, off, 64);
      const uint32_t otherKey = __shfl_down(key, off, 64);
      if (__lane_id() + off < 64 && otherKey == key) runMax = fmaxf(runMax, other);
    }
    if (valid && head && runMax > 0.0f && key < c.G - 1u && c.G > 0u)
      atomicMax(&L.cellDisp[c.keyBase + key], __float_as_uint(runMax));
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      disp = fmaxf(disp, __shfl_xor(disp, off, 64));
      move = fmaxf(move, __shfl_xor(move, off, 64));
    }
    // D and D_j: the workgroup's maxima, then one atomic each, and only when
    // it raises the value (they only grow during the pass, so a value read
    // here is a lower bound)
    if (__lane_id() == 0) {
      waveMax[threadIdx.x >> 6] = disp;
      waveMove[threadIdx.x >> 6] = move;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      float m = 0.0f, mv = 0.0f;
#pragma unroll
      for (int w = 0; w < BS / 64; ++w) {
        m = fmaxf(m, waveMax[w]);
        mv = fmaxf(mv, waveMove[w]);
      }
      const uint32_t bits = __float_as_uint(m);
      if (m > 0.0f && __hip_atomic_load(L.maxDisp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < bits)
        atomicMax(L.maxDisp, bits);
      uint32_t* const dj = L.maxDisp + 4 + min(passIdx, kPassSlots - 1);
      const uint32_t mbits = __float_as_uint(mv);
      if (mv > 0.0f && __hip_atomic_load(dj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < mbits)
        atomicMax(dj, mbits);
    }
    __syncthreads();
  });
}

// sph_finalise (oclsph_kernel.h:404-416). Also sets P.w to the fluid mass (0
// for obstacles) for the density field, and moves the diffused colour into V.w.
template <int BS>
__global__ __launch_bounds__(BS) void k_finalise(Work W, const uint8_t* __restrict__ T,
                                                 float4* __restrict__ P, float4* __restrict__ X,
                                                 float4* __restrict__ V,
                                                 const uint32_t* __restrict__ diffused,
                                                 int useDiffused) {
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    const uint32_t a = ch.begin + threadIdx.x;
    if (a >= ch.end) return;
    float4 p = P[a];
    if (T[a] == LF_OBSTACLE) {
      p.w = 0.0f;
      P[a] = p;
      return;
    }
    const float4 x = X[a];
    const float4 v = V[a];
    const float dX = p.x - x.x / c.scale;
    const float dY = p.y - x.y / c.scale;
    const float dZ = p.z - x.z / c.scale;
    X[a] = make_float4(p.x * c.scale, p.y * c.scale, p.z * c.scale, x.w);
    const float col = useDiffused ? __uint_as_float(diffused[a]) : v.w;
    V[a] = make_float4(__builtin_fmaf(dX, c.invDt, v.x) * c.vd,
                       __builtin_fmaf(dY, c.invDt, v.y) * c.vd,
                       __builtin_fmaf(dZ, c.invDt, v.z) * c.vd, col);
  });
}

__device__ inline void unpackARGB(uint32_t c, float out[4]) {
  out[0] = float((c >> 24) & 0xFFu);
  out[1] = float((c >> 16) & 0xFFu);
  out[2] = float((c >> 8) & 0xFFu);
  out[3] = float(c & 0xFFu);
}
// convert_uchar of a value already clamped to [8, 255]: round toward zero
__device__ inline uint32_t toU8(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 255.0f) return 255u;
  return uint32_t(v);
}

// sph_diffuse (oclsph_kernel.h:281-303): mean colour of the non-obstacle
// candidates, mix(c_a, mean * 1.33, dt/750), clamp [8, 255], truncate.
template <int BS>
__global__ __launch_bounds__(BS) void k_diffuse(Work W, const uint32_t* __restrict__ keys,
                                                const uint32_t* __restrict__ table,
                                                const uint8_t* __restrict__ T,
                                                const float4* __restrict__ V,
                                                uint32_t* __restrict__ diffused) {
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    const uint32_t a = ch.begin + threadIdx.x;
    if (a >= ch.end) return;
    const uint32_t ca = __float_as_uint(V[a].w);
    if (T[a] == LF_OBSTACLE) {
      diffused[a] = ca;
      return;
    }
    float mix[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int N = 0;
    forEachNeighbourCell(keys[a], table + c.keyBase, c.G, [&](uint32_t s, uint32_t e) {
      for (uint32_t b = s; b < e; ++b) {
        if (T[b] == LF_OBSTACLE) continue;
        float cb[4];
        unpackARGB(__float_as_uint(V[b].w), cb);
#pragma unroll
        for (int k = 0; k < 4; ++k) mix[k] += cb[k];
        ++N;
      }
    });
    float cf[4];
    unpackARGB(ca, cf);
    const float t = c.dt / 750.0f;
    uint32_t out = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const float y = (mix[k] / float(N)) * 1.33f;
      float o = __builtin_fmaf(y - cf[k], t, cf[k]);
      o = fminf(fmaxf(o, 8.0f), 255.0f);
      out = (out << 8) | toU8(o);
    }
    diffused[a] = out;
  });
}

// XSPH viscosity (absent in the reference): v_a += (c/rho0) sum_b m_b (v_b -
// v_a) W6(r) over b within h, Jacobi (Vin -> Vout), with m_b = P.w, the fluid
// mass after finalise (0 for obstacles).
template <int BS>
__global__ __launch_bounds__(BS) void k_xsph(Work W, const uint32_t* __restrict__ keys,
                                             const uint32_t* __restrict__ table,
                                             const uint8_t* __restrict__ T,
                                             const float4* __restrict__ P,
                                             const float4* __restrict__ Vin, Lists L,
                                             float4* __restrict__ Vout,
                                             const uint32_t* __restrict__ health, int passIdx) {
  // a diverged state: velocities copied through (Vout becomes the current V)
  const bool halted = (health[0] | health[1]) != 0u;
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    const uint32_t a = ch.begin + threadIdx.x;
    if (a >= ch.end) return;
    const float4 va = Vin[a];
    if (halted || T[a] == LF_OBSTACLE || c.xsph == 0.0f) {
      Vout[a] = va;
      return;
    }
    const float4 pa = P[a];
    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    auto term = [&](const float4& pb, const float4& vb, float d2) {
      // P.w is the fluid mass, 0 for obstacles: they add zero-weight terms.
      // W6 on d2 (no square root).
      const float t = c.hh - d2;
      const float mw = pb.w * (c.poly6 * ((t * t) * t));
      ax = __builtin_fmaf(mw, vb.x - va.x, ax);
      ay = __builtin_fmaf(mw, vb.y - va.y, ay);
      az = __builtin_fmaf(mw, vb.z - va.z, az);
    };
#if LF_XSPH_SPLIT
    // List walk in two gathers per block: positions, then the velocities of
    // the entries within h only (about half the list lies in the skin).
    forEachAcceptedSplit<kBlockXsph>(
        c, L, 3, passIdx, a, pa, keys[a], table + c.keyBase, P, Vin,
        [&](const float4& pb, const float4& vb, float d2) { term(pb, vb, d2); });
#else
    forEachAccepted<BS, kBlockXsph, false>(
        c, L, 3, false, passIdx, a, pa, keys[a], table + c.keyBase,
        [&](uint32_t b) { return PosVel{at(P, b), at(Vin, b)}; }, RelF32{pa},
        [&](uint32_t, const PosVel& r, float, float, float, float d2) { term(r.p, r.v, d2); },
        nullptr, 0, [&](uint32_t b) { return PosVel{at(P, b), at(Vin, b)}; });
#endif
    Vout[a] = make_float4(__builtin_fmaf(c.xsphRho, ax, va.x),
                          __builtin_fmaf(c.xsphRho, ay, va.y),
                          __builtin_fmaf(c.xsphRho, az, va.z), va.w);
  });
}

Lists listsOf(Ctx* c) {
  return Lists{c->nbr.p, c->nbrCount.p, c->P0.p, c->maxDisp.p, c->cellDisp.p, c->listEpoch.p,
               c->listR.p};
}

// The work of a pass that stops once the state has diverged (health counters).
Work haltingWork(Ctx* c, int queueSlot, bool skipDone = false) {
  Work W = workOf(c, queueSlot, skipDone);
  W.halt = c->health.p;
  return W;
}

}  // namespace

int pbfBuildLists(Ctx* c) {
  if (c->n == 0) return LF_OK;
  c->listFresh = true;
  c->passIdx = 1;
  const Work W = haltingWork(c, kSlotList);
  return LF_DISPATCH_BS(c, "pbf_build_list", k_build_list, W,
                        (const uint32_t*)c->localKey.p, (const uint32_t*)c->table.p,
                        (const uint8_t*)c->T[c->cur].p, listsOf(c),
                        int(c->n > size_t(kNbrBias) + 1));
}

// Calls f with std::integral_constant arguments for the workgroup size and
// the given flags, so one expression instantiates every kernel variant.
template <typename F> int withBS(const Ctx* c, F&& f) {
  switch (c->chunkThreads) {
    case 128: return f(std::integral_constant<int, 128>{});
    case 512: return f(std::integral_constant<int, 512>{});
    default: return f(std::integral_constant<int, 256>{});
  }
}
template <typename F> int withFlag(bool on, F&& f) {
  return on ? f(std::true_type{}) : f(std::false_type{});
}
// The pair terms' arithmetic: the hardware approximations under
// LF_FLAG_FAST_MATH, else the short correctly rounded forms where every
// scene's h allows them, else the compiler's IEEE sequences.
template <typename F> int withArith(const Ctx* c, bool fast, F&& f) {
  if (fast) return f(std::integral_constant<int, kArithHw>{});
  if (c->pairRn) return f(std::integral_constant<int, kArithRn>{});
  return f(std::integral_constant<int, kArithIeee>{});
}

// The packed records exist only when the grid was built with the flag.
static int requirePacked(Ctx* c, bool packed) {
  if (packed && !c->gridPacked)
    return fail(c, LF_ERR_STATE, "LF_FLAG_PACKED16: build the grid with the flag set");
  return LF_OK;
}

int pbfLambda(Ctx* c, bool skipDone) {
  if (c->n == 0) return LF_OK;
  if (int rc = requirePacked(c, (c->consts.flags & LF_FLAG_PACKED16) != 0)) return rc;
  const Work W = haltingWork(c, kSlotLambda, skipDone);
  const int fresh = int(c->listFresh);
  const uint32_t flags = c->consts.flags;
  const char* name = (flags & LF_FLAG_KAHAN)
                         ? ((flags & LF_FLAG_FAST_MATH) ? "pbf_lambda_kahan_fast" : "pbf_lambda_kahan")
                         : ((flags & LF_FLAG_FAST_MATH) ? "pbf_lambda_fast" : "pbf_lambda");
  const bool packed = (flags & LF_FLAG_PACKED16) != 0;
  return withBS(c, [&](auto bs) {
    return withFlag(flags & LF_FLAG_KAHAN, [&](auto kahan) {
      return withArith(c, flags & LF_FLAG_FAST_MATH, [&](auto ar) {
        return withFlag(packed, [&](auto pk) {
#if LF_PAIR_TILES
          if constexpr (!decltype(kahan)::value && decltype(ar)::value != kArithHw &&
                        !decltype(pk)::value)
            return launch(c, name, k_lambda<decltype(bs)::value, false, decltype(ar)::value, false, true>,
                          gridOf(c, W), dim3(decltype(bs)::value), 0, W,
                          (const uint32_t*)c->localKey.p, (const uint32_t*)c->table.p,
                          (const uint8_t*)c->T[c->cur].p, (const float4*)c->P.p, c->PL.p,
                          c->rho.p, listsOf(c), fresh, c->passIdx, 0,
                          int(c->n > size_t(kNbrBias) + 1), (const uint2*)c->Q.p, c->QL.p,
                          c->LAM.p);
#endif
          const int rebuild = int(c->passIdx > 1);
          const size_t ring =
              rebuild ? 2u * kNbrGroup * uint32_t(decltype(bs)::value) * sizeof(uint32_t) : 0u;
          return launch(c, name, k_lambda<decltype(bs)::value, decltype(kahan)::value,
                                          decltype(ar)::value, decltype(pk)::value>,
                        gridOf(c, W), dim3(decltype(bs)::value), ring, W,
                        (const uint32_t*)c->localKey.p, (const uint32_t*)c->table.p,
                        (const uint8_t*)c->T[c->cur].p, (const float4*)c->P.p, c->PL.p,
                        c->rho.p, listsOf(c), fresh, c->passIdx, rebuild,
                        int(c->n > size_t(kNbrBias) + 1), (const uint2*)c->Q.p, c->QL.p,
                        c->LAM.p);
        });
      });
    });
  });
}

int pbfDelta(Ctx* c, bool skipDone, bool final) {
  if (c->n == 0) return LF_OK;
  if (int rc = requirePacked(c, (c->consts.flags & LF_FLAG_PACKED16) != 0)) return rc;
  int grc = prepareColliderGrid(c);
  if (grc) return grc;
  if ((grc = waitPairReaders(c))) return grc;
  const Work W = haltingWork(c, kSlotDelta, skipDone);
  const bool fastMath = (c->consts.flags & LF_FLAG_FAST_MATH) != 0;
  const char* name = final ? (fastMath ? "pbf_delta_final_fast" : "pbf_delta_final")
                           : (fastMath ? "pbf_delta_fast" : "pbf_delta");
  const bool packed = (c->consts.flags & LF_FLAG_PACKED16) != 0;
  const int rc = withBS(c, [&](auto bs) {
    return withArith(c, fastMath, [&](auto ar) {
      return withFlag(final, [&](auto fin) {
        return withFlag(packed, [&](auto pk) {
#if LF_PAIR_TILES
          if constexpr (decltype(ar)::value != kArithHw && !decltype(pk)::value)
            return launch(c, name,
                          k_delta<decltype(bs)::value, decltype(ar)::value, decltype(fin)::value,
                                  false, true>,
                          gridOf(c, W), dim3(decltype(bs)::value), 0, W,
                          (const uint32_t*)c->localKey.p, (const uint32_t*)c->table.p,
                          (const uint8_t*)c->T[c->cur].p, (const float4*)c->PL.p, c->X[c->cur].p,
                          listsOf(c), int(c->listFresh), c->P.p, c->V[c->cur].p, c->colliderGrid,
                          (const uint32_t*)c->diffused.p,
                          int((c->consts.flags & LF_FLAG_DIFFUSE) != 0), (const uint2*)c->QL.p,
                          (const float*)c->LAM.p, c->Q.p, c->passIdx);
#endif
          return launch(c, name,
                        k_delta<decltype(bs)::value, decltype(ar)::value, decltype(fin)::value,
                                decltype(pk)::value>,
                        gridOf(c, W), dim3(decltype(bs)::value), 0, W,
                        (const uint32_t*)c->localKey.p, (const uint32_t*)c->table.p,
                        (const uint8_t*)c->T[c->cur].p, (const float4*)c->PL.p, c->X[c->cur].p,
                        listsOf(c), int(c->listFresh), c->P.p, c->V[c->cur].p, c->colliderGrid,
                        (const uint32_t*)c->diffused.p,
                        int((c->consts.flags & LF_FLAG_DIFFUSE) != 0), (const uint2*)c->QL.p,
                        (const float*)c->LAM.p, c->Q.p, c->passIdx);
        });
      });
    });
  });
  c->listFresh = false;  // p* now differs from the predicted positions
  ++c->passIdx;
  if (rc == LF_OK) notePositionsWritten(c);
  return rc;
}

int pbfFinalise(Ctx* c) {
  if (c->n == 0) return LF_OK;
  int rc0 = waitPairReaders(c);
  if (rc0) return rc0;
  const Work W = haltingWork(c, -1);
  const int rc = LF_DISPATCH_BS(c, "pbf_finalise", k_finalise, W, (const uint8_t*)c->T[c->cur].p,
                                c->P.p, c->X[c->cur].p, c->V[c->cur].p,
                                (const uint32_t*)c->diffused.p,
                                int((c->consts.flags & LF_FLAG_DIFFUSE) != 0));
  if (rc == LF_OK) notePositionsWritten(c);  // P.w
  return rc;
}

int pbfXsph(Ctx* c) {
  if (c->n == 0) return LF_OK;
  bool any = false;
  for (const PbfConsts& k : c->sceneConsts) any = any || k.xsph != 0.0f;
  if (!any) return LF_OK;
  const int nxt = 1 - c->cur;
  int rc = ensure(c, c->V[nxt], c->n, "velocities");
  if (rc) return rc;
  const Work W = workOf(c, kSlotXsph);
  rc = LF_DISPATCH_BS(c, "pbf_xsph", k_xsph, W, (const uint32_t*)c->localKey.p,
                      (const uint32_t*)c->table.p, (const uint8_t*)c->T[c->cur].p,
                      (const float4*)c->P.p, (const float4*)c->V[c->cur].p, listsOf(c),
                      c->V[nxt].p, (const uint32_t*)c->health.p, c->passIdx);
  if (rc) return rc;
  std::swap(c->V[0], c->V[1]);  // the XSPH output becomes the current V
  return LF_OK;
}

int pbfDiffuse(Ctx* c) {
  if (c->n == 0) return LF_OK;
  const Work W = haltingWork(c, -1);
  return LF_DISPATCH_BS(c, "pbf_diffuse", k_diffuse, W, (const uint32_t*)c->localKey.p,
                        (const uint32_t*)c->table.p, (const uint8_t*)c->T[c->cur].p,
                        (const float4*)c->V[c->cur].p, c->diffused.p);
}

}  // namespace lf
