// Neighbour-search grid on gfx950: predict + Morton key, and the stable sort
// and the dense cell-start table as one counting sort over the table's cells,
// with the payload reorder.
//
// Replaces the host pre-pass of ocl::SphSolver::advance: advectAndCopy
// (oclsph.hpp:289-325), computeBoundAndZindex (:327-387), the sort (:733-743)
// and the grid table (:746-767), so the whole step stays on the device.
#include <algorithm>

#include "lf_internal.hpp"

namespace lf {
namespace {

constexpr int kThreads = 256;

// ---- the grid build: a counting sort over the dense cell table ------------
// Every particle's sort key is its bucket: the scene's table base plus its
// Morton cell key, clamped to G (keys >= G belong to no cell, oclsph_kernel.h:
// 186-187, and share one bucket per scene). The buckets are the cells of the
// dense table, so the sort and the table come from one counting sort:
//   A k_predict_count  p*, key, count[key] += 1 (the old count is a
//                      provisional rank within the bucket)
//   B k_bucket_reduce  the sum of each tile of kBucketTile bucket counts
//     k_bucket_scan    exclusive scan of the counts = the table (first sorted
//                      index of every bucket, oclsph.hpp:758-765), each tile
//                      after the sum of the earlier tiles' sums
//   C k_bucket_scatter slot = table[key] + provisional rank; clears the counts
//   E k_sort_big       buckets of more than kFixupMax particles: their
//                      indices sorted within the bucket (one workgroup each)
//   D k_reorder        final slot = table[key] + the number of the bucket's
//                      particles with a smaller pre-step index (the stable
//                      order of std::stable_sort by key); gathers the state
// The provisional ranks come from atomics, so their order within a bucket is
// arbitrary; D (or E) restores the stable order, so the result is
// deterministic and equals a stable sort by (bucket, pre-step index).
constexpr uint32_t kFixupMax = 64;                       // D orders buckets up to this size
constexpr uint32_t kBucketItems = 16;                    // buckets per thread in the scan
constexpr uint32_t kBucketTile = kThreads * kBucketItems;  // 4096 buckets per workgroup
constexpr uint32_t kBigHeader = 3;  // big-bucket list: [0] count, [1] next, [2] workgroups done
constexpr unsigned kBigBlocks = 32;  // workgroups of k_sort_big (large buckets are rare)

// A: p* and its bucket per particle (state order), the bucket counts, and the
// count of particles whose p* is not finite (health[0]). Lanes of a wave that
// continue a run of equal buckets (the input is the last step's sorted order,
// so most do) take consecutive ranks from one atomic of the run's first lane.
template <int BS>
__global__ __launch_bounds__(BS) void k_predict_count(Work W, const float4* __restrict__ X,
                                                      const float4* __restrict__ V,
                                                      const uint8_t* __restrict__ T,
                                                      const lf_well* __restrict__ wells,
                                                      int nWells, uint32_t* __restrict__ count,
                                                      uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ ranks,
                                                      uint32_t* __restrict__ health,
                                                      const uint32_t* __restrict__ prevKey,
                                                      uint32_t* __restrict__ cellDisp) {
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    const uint32_t i = ch.begin + threadIdx.x;
    const bool valid = i < ch.end;
    uint32_t key = 0xFFFFFFFFu, old = 0xFFFFFFFFu;
    if (valid) {
      float px, py, pz;
      predictPstar(c, X[i], V[i], T[i], wells, nWells, px, py, pz);
      if (!(isfinite(px) && isfinite(py) && isfinite(pz))) atomicAdd(&health[0], 1u);
      key = c.keyBase + min(cellKey(c, px, py, pz), c.G);
      old = prevKey[i];
    }
    const uint32_t lane = __lane_id();
    // The cells whose displacement the last step's delta passes raised are
    // the cells of its sorted particles, which are this step's input in the
    // same order: each run of them is cleared once, before this step's passes
    // raise them again. (Input that did not come from the last step, after
    // an upload or a drain, leaves some cells at old values: larger than
    // this step's, so the list tests only become stricter.) The keys are
    // bounded without wrapping: they may be any value on the first step.
    const uint32_t prevOld = __shfl_up(old, 1, 64);
    if (valid && (lane == 0 || prevOld != old) && old < c.G - 1u && c.G > 0u)
      cellDisp[c.keyBase + old] = 0u;
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool start = valid && (lane == 0 || prev != key);
    const uint64_t starts = __ballot(start);
    const uint64_t ends = starts | ~__ballot(valid);  // a run also ends at the first idle lane
    const uint64_t upTo = (uint64_t(2) << lane) - 1u;   // lanes 0..lane (all for lane 63)
    const uint32_t first = 63u - uint32_t(__builtin_clzll((starts & upTo) | 1u));
    const uint64_t after = ends & ~upTo;
    const uint32_t next = after ? uint32_t(__builtin_ctzll(after)) : 64u;
    uint32_t base = 0;
    if (start) base = atomicAdd(&count[key], next - lane);
    base = __shfl(base, int(first), 64);
    if (valid) {
      keys[i] = key;
      ranks[i] = base + (lane - first);
    }
  });
}

// B1: per-workgroup sums of kBucketTile bucket counts. (Adding them up with
// atomics in A instead put every atomic on the few memory channels of one
// small array: 27 -> 180 us for A at 1M.)
__global__ __launch_bounds__(kThreads) void k_bucket_reduce(const uint32_t* __restrict__ count,
                                                            uint32_t nb,
                                                            uint32_t* __restrict__ tileSum) {
  __shared__ uint32_t sm[kThreads / 64];
  const uint32_t t0 = blockIdx.x * kBucketTile;
  uint32_t s = 0;
  if (t0 + kBucketTile <= nb) {
    const uint4* q = reinterpret_cast<const uint4*>(count + t0);
#pragma unroll
    for (uint32_t j = 0; j < kBucketItems / 4; ++j) {
      const uint4 v = q[j * kThreads + threadIdx.x];
      s += (v.x + v.y) + (v.z + v.w);
    }
  } else {
    for (uint32_t b = t0 + threadIdx.x; b < nb; b += kThreads) s += count[b];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (__lane_id() == 0) sm[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) t += sm[w];
    tileSum[blockIdx.x] = t;
  }
}

// B: the table from the counts. A workgroup's offset is the sum of the
// earlier tiles' sums (each workgroup adds them up itself: a few KB of reads,
// no separate scan launch); thread j then scans buckets [16j, 16j + 16) of the
// tile after the sums of the workgroup's earlier threads; table[nb] = n.
__global__ __launch_bounds__(kThreads) void k_bucket_scan(const uint32_t* __restrict__ count,
                                                          uint32_t nb,
                                                          const uint32_t* __restrict__ tileSum,
                                                          uint32_t* __restrict__ table) {
  __shared__ uint32_t sm[kThreads / 64];
  __shared__ uint32_t tileOff;
  const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
  {
    uint32_t e = 0;
    for (uint32_t t = threadIdx.x; t < blockIdx.x; t += kThreads) e += tileSum[t];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, 64);
    if (lane == 0u) sm[w] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
#pragma unroll
      for (int k = 0; k < kThreads / 64; ++k) t += sm[k];
      tileOff = t;
      // the last workgroup's offset plus its own sum is n
      if (blockIdx.x == gridDim.x - 1u) table[nb] = t + tileSum[blockIdx.x];
    }
    __syncthreads();
  }
  const uint32_t b0 = blockIdx.x * kBucketTile + threadIdx.x * kBucketItems;
  const bool whole = b0 + kBucketItems <= nb;
  uint32_t v[kBucketItems];
  if (whole) {
    const uint4* q = reinterpret_cast<const uint4*>(count + b0);
#pragma unroll
    for (uint32_t j = 0; j < kBucketItems / 4; ++j) {
      const uint4 x = q[j];
      v[4 * j] = x.x;
      v[4 * j + 1] = x.y;
      v[4 * j + 2] = x.z;
      v[4 * j + 3] = x.w;
    }
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kBucketItems; ++j) v[j] = b0 + j < nb ? count[b0 + j] : 0u;
  }
  uint32_t s = 0;
#pragma unroll
  for (uint32_t j = 0; j < kBucketItems; ++j) s += v[j];
  // exclusive prefix of s over the workgroup
  uint32_t x = s;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= uint32_t(o)) x += y;
  }
  if (lane == 63u) sm[w] = x;
  __syncthreads();
  uint32_t run = tileOff + x - s;
  for (uint32_t k = 0; k < w; ++k) run += sm[k];
  uint32_t t[kBucketItems];
#pragma unroll
  for (uint32_t j = 0; j < kBucketItems; ++j) {
    t[j] = run;
    run += v[j];
  }
  if (whole) {
    uint4* qt = reinterpret_cast<uint4*>(table + b0);
#pragma unroll
    for (uint32_t j = 0; j < kBucketItems / 4; ++j)
      qt[j] = make_uint4(t[4 * j], t[4 * j + 1], t[4 * j + 2], t[4 * j + 3]);
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kBucketItems; ++j)
      if (b0 + j < nb) table[b0 + j] = t[j];
  }
}

// C: each particle to its provisional slot. The first particle of a bucket
// (provisional rank 0) clears the bucket's count for the next step (only
// buckets holding particles were counted), lists the bucket for E when it
// holds more than kFixupMax particles, and counts it as overfull when it is
// a cell of the table holding more than `guard` particles (guard 0: no limit).
template <int BS>
__global__ __launch_bounds__(BS) void k_bucket_scatter(Work W, uint32_t n,
                                                       uint32_t* __restrict__ count,
                                                       const uint32_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ ranks,
                                                       const uint32_t* __restrict__ table,
                                                       uint32_t* __restrict__ tmpKey,
                                                       uint32_t* __restrict__ tmpVal,
                                                       uint32_t* __restrict__ big, uint32_t bigCap,
                                                       uint32_t guard, uint32_t* __restrict__ health) {
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    const uint32_t i = ch.begin + threadIdx.x;
    if (i >= ch.end) return;
    const uint32_t k = keys[i], r = ranks[i];
    const uint32_t s = table[k];
    const uint32_t pos = s + r;
    if (pos < n) {  // always, for consistent counts; guards the writes
      tmpKey[pos] = k;
      tmpVal[pos] = i;
    }
    if (r == 0u) {
      count[k] = 0u;
      const uint32_t len = table[k + 1u] - s;
      if (len > kFixupMax) {
        const uint32_t j = atomicAdd(&big[0], 1u);  // below n / (kFixupMax + 1) < bigCap
        if (j < bigCap) big[kBigHeader + j] = k;
      }
      if (guard && len > guard && k - c.keyBase < c.G) {
        atomicAdd(&health[1], 1u);
        atomicMax(&health[2], len);
      }
    }
  });
}

__device__ inline uint64_t laneMaskLt() { return (uint64_t(1) << __lane_id()) - 1u; }

// E: the buckets listed by C, one workgroup each (taken from a queue),
// sorted by particle index with a stable LSD radix sort of 8-bit digits over
// `passes` passes (even, so the result ends where it started), ping-ponging
// with the same range of `scratch`. Indices are distinct, so the order is the
// stable one. The last workgroup to leave clears the list for the next step.
constexpr int kBigThreads = 256;  // small, so that the (usually empty) launch finds room beside other work
constexpr int kBigWaves = kBigThreads / 64;
__global__ __launch_bounds__(kBigThreads) void k_sort_big(uint32_t* __restrict__ big,
                                                          const uint32_t* __restrict__ table,
                                                          uint32_t* __restrict__ vals,
                                                          uint32_t* __restrict__ scratch,
                                                          int passes) {
  __shared__ uint32_t bucket;
  __shared__ uint32_t digitBase[256];
  __shared__ uint32_t waveCnt[kBigWaves][256];
  const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
  // nothing listed (the usual case): no queue, nothing to clear
  if (threadIdx.x == 0) bucket = big[0];
  __syncthreads();
  if (bucket == 0u) return;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint32_t j = atomicAdd(&big[1], 1u);
      bucket = j < big[0] ? big[kBigHeader + j] : 0xFFFFFFFFu;
    }
    __syncthreads();
    const uint32_t k = bucket;
    if (k == 0xFFFFFFFFu) break;
    const uint32_t s0 = table[k], s1 = table[k + 1u];
    uint32_t* src = vals;
    uint32_t* dst = scratch;
    for (int p = 0; p < passes; ++p) {
      const uint32_t shift = 8u * uint32_t(p);
      if (threadIdx.x < 256) digitBase[threadIdx.x] = 0u;
      __syncthreads();
      for (uint32_t e = s0 + threadIdx.x; e < s1; e += kBigThreads)
        atomicAdd(&digitBase[(src[e] >> shift) & 255u], 1u);
      __syncthreads();
      if (w == 0) {  // exclusive scan of the 256 digit counts, 4 per lane
        uint32_t c4[4], t = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          c4[q] = digitBase[lane * 4u + q];
          t += c4[q];
        }
        uint32_t x = t;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t y = __shfl_up(x, o, 64);
          if (lane >= uint32_t(o)) x += y;
        }
        uint32_t run = s0 + x - t;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          digitBase[lane * 4u + q] = run;
          run += c4[q];
        }
      }
      __syncthreads();
      // rounds of kBigThreads values in input order: rank within the wave by
      // ballot, then per digit the earlier waves' counts
      for (uint32_t t0 = s0; t0 < s1; t0 += kBigThreads) {
        const uint32_t e = t0 + threadIdx.x;
        const bool valid = e < s1;
        const uint32_t v = valid ? src[e] : 0u;
        const uint32_t d = (v >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
          const uint32_t bit = (d >> b) & 1u;
          const uint64_t bb = __ballot(bit);
          peers &= bit ? bb : ~bb;
        }
        for (uint32_t q = threadIdx.x; q < kBigWaves * 256u; q += kBigThreads)
          (&waveCnt[0][0])[q] = 0u;
        __syncthreads();
        if (valid && lane == uint32_t(__builtin_ctzll(peers))) waveCnt[w][d] = uint32_t(__popcll(peers));
        __syncthreads();
        if (threadIdx.x < 256) {
          uint32_t run = digitBase[threadIdx.x];
          for (int x = 0; x < kBigWaves; ++x) {
            const uint32_t cnt = waveCnt[x][threadIdx.x];
            waveCnt[x][threadIdx.x] = run;
            run += cnt;
          }
          digitBase[threadIdx.x] = run;
        }
        __syncthreads();
        if (valid) dst[waveCnt[w][d] + uint32_t(__popcll(peers & laneMaskLt()))] = v;
        __syncthreads();
      }
      uint32_t* tmp = src;
      src = dst;
      dst = tmp;
    }
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&big[2], 1u) == gridDim.x - 1u) {
      big[0] = 0u;
      big[1] = 0u;
      big[2] = 0u;
      __threadfence();
    }
  }
}

// D: the final slot of each particle, and the state gathered into sorted
// order; p* is recomputed from the gathered x, v (same arithmetic as A, so
// identical bits) instead of being stored by A and gathered, and with it the
// scene-local cell key the pair passes use and P0, p* clamped to the bounds,
// from which the neighbour lists are built. v stays the pre-gravity velocity:
// the reference hands the device the velocity copied before the gravity
// update (oclsph.hpp:300,831).
template <int BS>
__global__ __launch_bounds__(BS) void k_reorder(
    Work W, uint32_t n, const uint32_t* __restrict__ tmpKey, const uint32_t* __restrict__ tmpVal,
    const uint32_t* __restrict__ table, const float4* __restrict__ Xin,
    const float4* __restrict__ Vin, const uint8_t* __restrict__ Tin,
    const uint32_t* __restrict__ IDin, const lf_well* __restrict__ wells, int nWells,
    float4* __restrict__ Xout, float4* __restrict__ Vout, uint8_t* __restrict__ Tout,
    uint32_t* __restrict__ IDout, float4* __restrict__ P, uint32_t* __restrict__ localKey,
    float4* __restrict__ P0, uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut,
    const uint32_t* __restrict__ health, uint32_t* __restrict__ healthHost,
    uint2* __restrict__ Q) {
  // the step's health counts are final (A and C have run): their copy in
  // pinned host memory, read by the host without waiting for the stream
  if (blockIdx.x == 0 && threadIdx.x < 3 && healthHost) healthHost[threadIdx.x] = health[threadIdx.x];
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    const uint32_t s = ch.begin + threadIdx.x;
    if (s >= ch.end) return;
    const uint32_t k = tmpKey[s];
    const uint32_t i = min(tmpVal[s], n - 1u);  // a permutation; the clamp only bounds reads
    // the state gathers are issued before the rank search (they do not depend on it)
    const float4 x = Xin[i];
    const float4 v = Vin[i];
    const uint8_t t = Tin[i];
    const uint32_t id = IDin[i];
    const uint32_t b0 = table[k], b1 = table[k + 1u];
    uint32_t f = s;  // buckets above kFixupMax were put in order by k_sort_big
    if (b1 - b0 <= kFixupMax) {
      // the bucket's pre-step indices 8 at a time (indices clamped into the
      // bucket): one memory latency per 8 entries, not per entry
      f = b0;
      for (uint32_t u0 = b0; u0 < b1; u0 += 8u) {
        uint32_t w[8];
#pragma unroll
        for (uint32_t u = 0; u < 8u; ++u) w[u] = tmpVal[min(u0 + u, b1 - 1u)];
#pragma unroll
        for (uint32_t u = 0; u < 8u; ++u) f += (u0 + u < b1 && w[u] < i) ? 1u : 0u;
      }
    }
    f = min(f, n - 1u);
    float px, py, pz;
    predictPstar(c, x, v, t, wells, nWells, px, py, pz);
    Xout[f] = x;
    Vout[f] = v;
    Tout[f] = t;
    IDout[f] = id;
    P[f] = make_float4(px, py, pz, x.w);
    localKey[f] = cellKey(c, px, py, pz);
    P0[f] = make_float4(clampToBounds(c, 0, px), clampToBounds(c, 1, py), clampToBounds(c, 2, pz),
                        x.w);
    keysOut[f] = k;
    valsOut[f] = i;
    if (Q) Q[f] = packPos(c, px, py, pz);  // LF_FLAG_PACKED16 records
  });
}

// Plain streaming copy, 16 B per lane per iteration (bandwidth probe).
__global__ __launch_bounds__(kThreads) void k_copy_probe(const float4* __restrict__ src,
                                                         float4* __restrict__ dst, size_t n) {
  const size_t stride = size_t(gridDim.x) * kThreads;
  for (size_t i = size_t(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// ---- drains: stable erase of drained fluid particles --------------------------
constexpr uint32_t kCompactPerThread = 4;
constexpr uint32_t kCompactBlock = kThreads * kCompactPerThread;

// oclsph.hpp:693-706: a fluid particle is erased when glm::distance(centre,
// position) < width for some drain (world units; dot product left to right).
__device__ inline bool drained(const float4 x, uint8_t t, const lf_drain* __restrict__ d, int nd) {
  if (t == LF_OBSTACLE) return false;
  for (int i = 0; i < nd; ++i) {
    const float dx = d[i].centre[0] - x.x, dy = d[i].centre[1] - x.y, dz = d[i].centre[2] - x.z;
    if (__builtin_sqrtf((dx * dx + dy * dy) + dz * dz) < d[i].width) return true;
  }
  return false;
}

__global__ __launch_bounds__(kThreads) void k_drain_count(uint32_t n, const float4* __restrict__ X,
                                                          const uint8_t* __restrict__ T,
                                                          const lf_drain* __restrict__ d, int nd,
                                                          uint32_t* __restrict__ blockCount) {
  __shared__ uint32_t sm[kThreads / 64];
  uint32_t keep = 0;
#pragma unroll
  for (uint32_t j = 0; j < kCompactPerThread; ++j) {
    const uint32_t i = blockIdx.x * kCompactBlock + j * kThreads + threadIdx.x;
    if (i < n && !drained(X[i], T[i], d, nd)) ++keep;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) keep += __shfl_xor(keep, o, 64);
  if (__lane_id() == 0) sm[threadIdx.x >> 6] = keep;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < kThreads / 64; ++w) t += sm[w];
    blockCount[blockIdx.x] = t;
  }
}

// Exclusive scan of the block counts in place (one workgroup of 1024), total
// into counts[nb].
__global__ __launch_bounds__(1024) void k_scan_counts(uint32_t* __restrict__ counts, uint32_t nb) {
  __shared__ uint32_t sm[16];
  const uint32_t per = (nb + 1023) / 1024;
  const uint32_t lo = threadIdx.x * per, hi = min(lo + per, nb);
  uint32_t s = 0;
#pragma unroll 8
  for (uint32_t i = lo; i < hi; ++i) s += counts[i];
  uint32_t x = s;
  const int lane = __lane_id(), w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sm[w] = x;
  __syncthreads();
  uint32_t wb = 0, tot = 0;
  for (int i = 0; i < 16; ++i) {
    if (i < w) wb += sm[i];
    tot += sm[i];
  }
  uint32_t run = wb + x - s;
  for (uint32_t i = lo; i < hi; ++i) {
    const uint32_t t = counts[i];
    counts[i] = run;
    run += t;
  }
  if (threadIdx.x == 0) counts[nb] = tot;
}

// Kept particles to their scanned slots, in order (ballot prefix per round).
__global__ __launch_bounds__(kThreads) void k_drain_scatter(
    uint32_t n, const float4* __restrict__ Xin, const float4* __restrict__ Vin,
    const uint8_t* __restrict__ Tin, const uint32_t* __restrict__ IDin,
    const lf_drain* __restrict__ d, int nd, const uint32_t* __restrict__ blockOffset,
    float4* __restrict__ Xout, float4* __restrict__ Vout, uint8_t* __restrict__ Tout,
    uint32_t* __restrict__ IDout) {
  __shared__ uint32_t sm[kThreads / 64];
  uint32_t run = blockOffset[blockIdx.x];
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  for (uint32_t j = 0; j < kCompactPerThread; ++j) {
    const uint32_t i = blockIdx.x * kCompactBlock + j * kThreads + threadIdx.x;
    float4 x = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint8_t t = 0;
    bool keep = false;
    if (i < n) {
      x = Xin[i];
      t = Tin[i];
      keep = !drained(x, t, d, nd);
    }
    const uint64_t bb = __ballot(keep);
    if (lane == 0) sm[w] = uint32_t(__popcll(bb));
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int k = 0; k < kThreads / 64; ++k) {
      if (k < w) before += sm[k];
      total += sm[k];
    }
    __syncthreads();
    if (keep) {
      const uint32_t o = run + before + uint32_t(__popcll(bb & ((uint64_t(1) << lane) - 1u)));
      Xout[o] = x;
      Vout[o] = Vin[i];
      Tout[o] = t;
      IDout[o] = IDin[i];
    }
    run += total;
  }
}

// ---- spatial queries (oclsph.hpp:770-791) ---------------------------------------
// Per query point (world units): the fluid particles of its own cell, their
// count and the mean of their ARGB bytes packed back (truncated), from the
// colours the step started with.
__global__ __launch_bounds__(kThreads) void k_queries(uint32_t nq, const float* __restrict__ pts,
                                                      uint32_t n, PbfConsts c,
                                                      const uint32_t* __restrict__ table,
                                                      const uint8_t* __restrict__ T,
                                                      const float4* __restrict__ V,
                                                      uint32_t* __restrict__ out) {
  const uint32_t q = blockIdx.x * kThreads + threadIdx.x;
  if (q >= nq) return;
  const uint32_t key = mortonEncode(cellCoord((pts[3 * q] / c.scale - c.minExt[0]) / c.h),
                                    cellCoord((pts[3 * q + 1] / c.scale - c.minExt[1]) / c.h),
                                    cellCoord((pts[3 * q + 2] / c.scale - c.minExt[2]) / c.h));
  uint32_t N = 0, packed = 0;
  if (n > 0 && key + 1u < c.G) {  // with no particles the table was not rebuilt
    float avg[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    const uint32_t e = table[key + 1u];
    for (uint32_t a = table[key]; a < e; ++a) {
      if (T[a] != LF_FLUID) continue;
      const uint32_t col = __float_as_uint(V[a].w);
      ++N;
#pragma unroll
      for (int k = 0; k < 4; ++k) avg[k] += float((col >> (24 - 8 * k)) & 0xFFu);
    }
    if (N)
#pragma unroll
      for (int k = 0; k < 4; ++k) packed = (packed << 8) | (uint32_t(uint8_t(avg[k] / float(N))) & 0xFFu);
  }
  out[2 * q] = N;
  out[2 * q + 1] = packed;
}

}  // namespace

int applyDrains(Ctx* c) {
  const size_t n = c->n;
  if (n == 0 || c->hostDrains.empty()) return LF_OK;
  const uint32_t nb = uint32_t((n + kCompactBlock - 1) / kCompactBlock);
  int rc;
  const int nxt = 1 - c->cur;
  if ((rc = ensure(c, c->compactBlock, size_t(nb) + 1, "drain counts"))) return rc;
  if ((rc = ensure(c, c->X[nxt], n, "positions"))) return rc;
  if ((rc = ensure(c, c->V[nxt], n, "velocities"))) return rc;
  if ((rc = ensure(c, c->T[nxt], n, "types"))) return rc;
  if ((rc = ensure(c, c->ID[nxt], n, "ids"))) return rc;
  const int nd = int(c->hostDrains.size());
  rc = launch(c, "drain_count", k_drain_count, dim3(nb), dim3(kThreads), 0, uint32_t(n),
              (const float4*)c->X[c->cur].p, (const uint8_t*)c->T[c->cur].p,
              (const lf_drain*)c->drains.p, nd, c->compactBlock.p);
  if (rc) return rc;
  rc = launch(c, "drain_scan", k_scan_counts, dim3(1), dim3(1024), 0, c->compactBlock.p, nb);
  if (rc) return rc;
  rc = launch(c, "drain_scatter", k_drain_scatter, dim3(nb), dim3(kThreads), 0, uint32_t(n),
              (const float4*)c->X[c->cur].p, (const float4*)c->V[c->cur].p,
              (const uint8_t*)c->T[c->cur].p, (const uint32_t*)c->ID[c->cur].p,
              (const lf_drain*)c->drains.p, nd, (const uint32_t*)c->compactBlock.p, c->X[nxt].p,
              c->V[nxt].p, c->T[nxt].p, c->ID[nxt].p);
  if (rc) return rc;
  uint32_t kept = 0;
  LF_HIP(c, hipMemcpyAsync(&kept, c->compactBlock.p + nb, 4, hipMemcpyDeviceToHost, c->stream));
  LF_HIP(c, hipStreamSynchronize(c->stream));
  c->cur = nxt;
  c->n = kept;
  c->sceneBegin = {0u, kept};
  return LF_OK;
}

int answerQueries(Ctx* c) {
  if (c->nQueries == 0) return LF_OK;
  int rc = ensure(c, c->queryOut, 2 * size_t(c->nQueries), "query results");
  if (rc) return rc;
  rc = launch(c, "queries", k_queries, dim3(blocksFor(c->nQueries, kThreads)), dim3(kThreads), 0,
              c->nQueries, (const float*)c->queryPts.p, uint32_t(c->n), c->consts,
              (const uint32_t*)c->table.p,
              (const uint8_t*)c->T[c->cur].p, (const float4*)c->V[c->cur].p, c->queryOut.p);
  if (rc) return rc;
  c->queriesAnswered = true;
  return LF_OK;
}

int copyProbe(Ctx* c, size_t bytes, int reps, float* ms) {
  const size_t n = bytes / 16;
  float4 *a = nullptr, *b = nullptr;
  LF_HIP(c, hipMalloc(reinterpret_cast<void**>(&a), n * 16));
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&b), n * 16);
  if (e != hipSuccess) {
    (void)hipFree(a);
    return hipFail(c, e, "copy probe");
  }
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  int rc = LF_OK;
  (void)hipMemsetAsync(a, 0, n * 16, c->stream);
  const unsigned grid = unsigned(std::min<size_t>((n + kThreads - 1) / kThreads, 256 * 64));
  rc = launch(c, "copy_probe", k_copy_probe, dim3(grid), dim3(kThreads), 0, (const float4*)a, b, n);
  (void)hipEventRecord(t0, c->stream);
  for (int r = 0; r < reps && rc == LF_OK; ++r)
    rc = launch(c, "copy_probe", k_copy_probe, dim3(grid), dim3(kThreads), 0, (const float4*)a, b,
                n);
  (void)hipEventRecord(t1, c->stream);
  (void)hipEventSynchronize(t1);
  float total = 0.0f;
  (void)hipEventElapsedTime(&total, t0, t1);
  *ms = total / float(reps);
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  (void)hipFree(a);
  (void)hipFree(b);
  return rc;
}

// Grows b to count elements, zero-filled when it grows (the bucket counts and
// the big-bucket list are kept zero between steps by the kernels that use them).
template <typename T> static int ensureZeroed(Ctx* c, DevBuf<T>& b, size_t count, const char* what) {
  const bool grows = count > b.cap;
  int rc = ensure(c, b, count, what);
  if (rc || !grows) return rc;
  LF_HIP(c, hipMemsetAsync(b.p, 0, b.cap * sizeof(T), c->stream));
  return LF_OK;
}

// The pinned host copy of the health counters (written by k_reorder).
static int ensureHealthHost(Ctx* c) {
  if (c->healthHost) return LF_OK;
  LF_HIP(c, hipHostMalloc(reinterpret_cast<void**>(&c->healthHost), 4 * sizeof(uint32_t)));
  std::memset(c->healthHost, 0, 4 * sizeof(uint32_t));
  return LF_OK;
}

int gridBuild(Ctx* c) {
  const size_t n = c->n;
  int rc;
  // This step writes the other P / table pair; the density field that last
  // read that pair (on the marching-cubes stream, two steps ago) comes first.
  std::swap(c->P, c->Palt);
  std::swap(c->table, c->tableAlt);
  std::swap(c->pairGuard, c->pairGuardAlt);
  if (c->pairGuard.pending) {
    c->pairGuard.pending = false;
    LF_HIP(c, hipStreamWaitEvent(c->stream, c->pairGuard.ev, 0));
  }
  if ((rc = prepareScenes(c))) return rc;
  const uint32_t nb = c->bucketTotal;
  const uint32_t nTiles = (nb + kBucketTile - 1) / kBucketTile;
  const int nxt = 1 - c->cur;
  if ((rc = ensure(c, c->sortKey[0], n, "sort keys"))) return rc;
  if ((rc = ensure(c, c->sortKey[1], n, "sort keys"))) return rc;
  if ((rc = ensure(c, c->sortVal[0], n, "sort values"))) return rc;
  if ((rc = ensure(c, c->sortVal[1], n, "sort values"))) return rc;
  if ((rc = ensure(c, c->X[nxt], n, "positions"))) return rc;
  if ((rc = ensure(c, c->V[nxt], n, "velocities"))) return rc;
  if ((rc = ensure(c, c->T[nxt], n, "types"))) return rc;
  if ((rc = ensure(c, c->ID[nxt], n, "ids"))) return rc;
  if ((rc = ensure(c, c->P, n, "p*"))) return rc;
  if ((rc = ensure(c, c->Palt, n, "p*"))) return rc;  // both pairs: no growth in a graph
  if ((rc = ensure(c, c->PL, n, "p*+lambda"))) return rc;
  if ((rc = ensureZeroed(c, c->localKey, n, "cell keys"))) return rc;
  if ((rc = ensure(c, c->nbr, nbrCapacity(n), "neighbour lists"))) return rc;
  if ((rc = ensure(c, c->nbrCount, n, "neighbour counts"))) return rc;
  if ((rc = ensure(c, c->P0, n, "list positions"))) return rc;
  if ((rc = ensure(c, c->maxDisp, 4 + 16, "displacement"))) return rc;
  if ((rc = ensure(c, c->listEpoch, n, "list epochs"))) return rc;
  if ((rc = ensure(c, c->listR, n, "list positions"))) return rc;
  if ((rc = ensure(c, c->rho, n, "density"))) return rc;
  if ((rc = ensure(c, c->diffused, n, "diffused"))) return rc;
  if ((rc = ensure(c, c->table, size_t(nb) + 1, "cell table"))) return rc;
  if ((rc = ensure(c, c->tableAlt, size_t(nb) + 1, "cell table"))) return rc;
  if ((rc = ensureZeroed(c, c->cellDisp, size_t(nb), "cell displacement"))) return rc;
  if ((rc = ensureZeroed(c, c->cellCount, size_t(nb), "bucket counts"))) return rc;
  if ((rc = ensure(c, c->bucketTile, size_t(nTiles), "bucket tiles"))) return rc;
  if ((rc = ensureHealthHost(c))) return rc;
  if ((rc = ensureZeroed(c, c->bigBuckets, kBigHeader + n / kFixupMax + 2, "large buckets")))
    return rc;
  if ((rc = ensureZeroed(c, c->health, 4, "health"))) return rc;
  const bool packed = (c->consts.flags & LF_FLAG_PACKED16) != 0;
  if (packed) {
    if ((rc = ensure(c, c->Q, n, "packed positions"))) return rc;
    if ((rc = ensure(c, c->QL, n, "packed positions"))) return rc;
    if ((rc = ensure(c, c->LAM, n, "lambda"))) return rc;
  }
  if (n == 0) return LF_OK;
  const Work W = workOf(c, -1);  // no halt: the grid is built whatever the state
  rc = LF_DISPATCH_BS(c, "predict_count", k_predict_count, W, (const float4*)c->X[c->cur].p,
                      (const float4*)c->V[c->cur].p, (const uint8_t*)c->T[c->cur].p,
                      (const lf_well*)c->wells.p, c->nWells, c->cellCount.p, c->sortKey[1].p,
                      c->sortVal[1].p, c->health.p, (const uint32_t*)c->localKey.p,
                      c->cellDisp.p);
  if (rc) return rc;
  rc = launch(c, "bucket_reduce", k_bucket_reduce, dim3(nTiles), dim3(kThreads), 0,
              (const uint32_t*)c->cellCount.p, nb, c->bucketTile.p);
  if (rc) return rc;
  rc = launch(c, "bucket_scan", k_bucket_scan, dim3(nTiles), dim3(kThreads), 0,
              (const uint32_t*)c->cellCount.p, nb, (const uint32_t*)c->bucketTile.p, c->table.p);
  if (rc) return rc;
  rc = LF_DISPATCH_BS(c, "bucket_scatter", k_bucket_scatter, W, uint32_t(n), c->cellCount.p,
                      (const uint32_t*)c->sortKey[1].p, (const uint32_t*)c->sortVal[1].p,
                      (const uint32_t*)c->table.p, c->sortKey[0].p, c->sortVal[0].p,
                      c->bigBuckets.p, uint32_t(c->bigBuckets.cap - kBigHeader), c->guardCell,
                      c->health.p);
  if (rc) return rc;
  // distinct indices below n: 8-bit digits, an even number of passes
  int bits = 1;
  while (bits < 32 && (size_t(1) << bits) < n) ++bits;
  const int passes = ((bits + 7) / 8 + 1) & ~1;
  rc = launch(c, "sort_big", k_sort_big, dim3(kBigBlocks), dim3(kBigThreads), 0, c->bigBuckets.p,
              (const uint32_t*)c->table.p, c->sortVal[0].p, c->sortVal[1].p, passes);
  if (rc) return rc;
  rc = LF_DISPATCH_BS(c, "reorder", k_reorder, W, uint32_t(n), (const uint32_t*)c->sortKey[0].p,
                      (const uint32_t*)c->sortVal[0].p, (const uint32_t*)c->table.p,
                      (const float4*)c->X[c->cur].p, (const float4*)c->V[c->cur].p,
                      (const uint8_t*)c->T[c->cur].p, (const uint32_t*)c->ID[c->cur].p,
                      (const lf_well*)c->wells.p, c->nWells, c->X[nxt].p, c->V[nxt].p,
                      c->T[nxt].p, c->ID[nxt].p, c->P.p, c->localKey.p, c->P0.p,
                      c->sortKey[1].p, c->sortVal[1].p, (const uint32_t*)c->health.p,
                      c->healthHost, packed ? c->Q.p : (uint2*)nullptr);
  if (rc) return rc;
  notePositionsWritten(c);  // P (reorder) and the table
  c->gridPacked = packed;
  c->keysIdx = 1;
  c->gridG = c->consts.G;
  c->gridTotal = c->tableTotal;
  c->cur = nxt;
  c->gridValid = true;
  return LF_OK;
}

}  // namespace lf
