// Stable LSD radix sort of (Morton key, particle index) pairs on gfx950.
//
// Replaces the host std::sort of oclsph.hpp:733-743. The reference sort is
// unstable; this one is stable, so ties keep the pre-sort order (the CPU
// restatement uses std::stable_sort for the same order).
//
// Keys are 30-bit Morton codes (10 bits per axis, curves.h:56-72): three
// passes of 10-bit digits. Each pass is three kernels over tiles of 2048
// pairs:
//   hist    per-tile digit counts (512 threads x 4, wave-aggregated LDS
//           counters), stored tile-major: one contiguous 4 KB row per tile
//   scan    per-digit exclusive scan over tiles (16 digits x 64 tile
//           segments per workgroup)
//   scatter digit bases + stable in-tile rank (256 threads x 8, per-wave
//           ballot match, LDS running counters in (wave, slot, lane) = input
//           order) -> output
#include "lf_internal.hpp"

namespace lf {
namespace {

constexpr int kDigitBits = 10;
constexpr uint32_t kRadix = 1u << kDigitBits;
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kItems = 8;
constexpr uint32_t kTile = kThreads * kItems;     // 2048 pairs
constexpr uint32_t kWaveTile = 64 * kItems;       // 512 pairs
constexpr int kPasses = 3;

__device__ inline uint64_t laneMaskLt() {
  return (uint64_t(1) << __lane_id()) - 1u;
}

// Lanes of the wave (among `active`) holding the same digit as this lane.
__device__ inline uint64_t matchDigit(uint32_t d, uint64_t active) {
  uint64_t m = active;
#pragma unroll
  for (int b = 0; b < kDigitBits; ++b) {
    const uint32_t bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(bit);
    m &= bit ? bb : ~bb;
  }
  return m;
}

// The histogram kernel covers a tile with 512 threads x 4 keys (14.3 us per
// pass at 1M against 15.5 with the scatter's 256 x 8).
constexpr int kHistThreads = 512;
constexpr int kHistItems = int(kTile) / kHistThreads;
__global__ __launch_bounds__(kHistThreads) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                             uint32_t n, int shift,
                                                             uint32_t* __restrict__ hist,
                                                             uint32_t nTiles) {
  __shared__ uint32_t h[kRadix];
  for (uint32_t i = threadIdx.x; i < kRadix; i += kHistThreads) h[i] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kTile;
#pragma unroll
  for (int j = 0; j < kHistItems; ++j) {
    const uint32_t e = base + j * kHistThreads + threadIdx.x;
    const bool valid = e < n;
    const uint32_t d = valid ? (keys[e] >> shift) & (kRadix - 1u) : 0u;
    const uint64_t peers = matchDigit(d, __ballot(valid));
    const int leader = peers ? __builtin_ctzll(peers) : -1;
    if (valid && __lane_id() == leader) atomicAdd(&h[d], uint32_t(__popcll(peers)));
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kRadix; i += kHistThreads) hist[size_t(blockIdx.x) * kRadix + i] = h[i];
}

// Block-wide exclusive scan of one value per thread (256 threads); returns
// the exclusive prefix, *total the block sum.
__device__ inline uint32_t blockExclusiveScan(uint32_t v, uint32_t* sm /*[kWaves]*/,
                                              uint32_t* total) {
  const int lane = __lane_id();
  const int w = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sm[w] = x;
  __syncthreads();
  uint32_t waveBase = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kWaves; ++i) {
    const uint32_t s = sm[i];
    if (i < w) waveBase += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return waveBase + x - v;
}

// Exclusive scan over tiles of each digit's counts, in place (hist[t][d]),
// digit totals into tot[d]. A workgroup of kScanThreads takes kScanDigits
// digits; its threads split the tiles into kScanSegs segments per digit (a few
// tiles each), so each load of a wave reads runs of kScanDigits digits.
// 16 digits (64-byte runs, one wave per digit's segments) measured at 1M:
// 11.0 us per pass against 12.2 (8 digits), 12.1 (32) and 15.1 (4).
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kScanDigits = 16;
constexpr uint32_t kScanSegs = kScanThreads / kScanDigits;
// the segment sums of a digit are scanned by kScanSegs consecutive lanes: a
// part of one wave, or whole waves
constexpr uint32_t kSegWidth = kScanSegs < 64 ? kScanSegs : 64;
static_assert(kRadix % kScanDigits == 0, "whole digit groups per workgroup");
static_assert(kScanSegs >= 64 ? kScanSegs % 64 == 0 : 64 % kScanSegs == 0, "segments tile waves");
__global__ __launch_bounds__(kScanThreads) void k_radix_scan(uint32_t* __restrict__ hist,
                                                             uint32_t nTiles,
                                                             uint32_t* __restrict__ tot) {
  __shared__ uint32_t segSum[kScanSegs][kScanDigits];
  __shared__ uint32_t waveTot[kScanThreads / 64];
  const uint32_t dl = threadIdx.x % kScanDigits, seg = threadIdx.x / kScanDigits;
  const uint32_t d = blockIdx.x * kScanDigits + dl;
  const uint32_t per = (nTiles + kScanSegs - 1) / kScanSegs;
  const uint32_t t0 = min(seg * per, nTiles), t1 = min(t0 + per, nTiles);
  uint32_t s = 0;
#pragma unroll 4
  for (uint32_t t = t0; t < t1; ++t) s += hist[size_t(t) * kRadix + d];
  segSum[seg][dl] = s;
  __syncthreads();
  {
    // exclusive prefix of the segment sums: thread j scans segment
    // j % kScanSegs of digit j / kScanSegs (a shuffle scan over kSegWidth
    // lanes, then the totals of the digit's earlier waves)
    const uint32_t sd = threadIdx.x / kScanSegs, ss = threadIdx.x % kScanSegs;
    const uint32_t lane = __lane_id(), w = threadIdx.x >> 6;
    const uint32_t v = segSum[ss][sd];
    uint32_t x = v;
#pragma unroll
    for (uint32_t o = 1; o < kSegWidth; o <<= 1) {
      const uint32_t y = __shfl_up(x, o, kSegWidth);
      if (lane % kSegWidth >= o) x += y;
    }
    uint32_t before = 0;
    if constexpr (kScanSegs > 64) {
      if (lane == 63u) waveTot[w] = x;
      __syncthreads();
      for (uint32_t i = (sd * kScanSegs) / 64; i < w; ++i) before += waveTot[i];
    }
    segSum[ss][sd] = before + x - v;
    if (ss == kScanSegs - 1) tot[blockIdx.x * kScanDigits + sd] = before + x;
  }
  __syncthreads();
  uint32_t run = segSum[seg][dl];
#pragma unroll 4
  for (uint32_t t = t0; t < t1; ++t) {
    const size_t i = size_t(t) * kRadix + d;
    const uint32_t v = hist[i];
    hist[i] = run;
    run += v;
  }
}

__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
    uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut, uint32_t n, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ tot) {
  __shared__ uint32_t cnt[kWaves][kRadix];  // per-wave counts -> wave offsets
  __shared__ uint32_t base[kRadix];         // this tile's global base per digit
  __shared__ uint32_t sm[kWaves];
  const int lane = __lane_id();
  const int w = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < kWaves * kRadix; i += kThreads) (&cnt[0][0])[i] = 0;

  // digit bases: exclusive scan of the digit totals, plus this tile's offset
  constexpr int kPer = kRadix / kThreads;  // 4 digits per thread
  uint32_t t4[kPer], s = 0;
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    t4[i] = tot[threadIdx.x * kPer + i];
    s += t4[i];
  }
  uint32_t total;
  uint32_t run = blockExclusiveScan(s, sm, &total);
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const uint32_t d = threadIdx.x * kPer + i;
    base[d] = run + hist[size_t(blockIdx.x) * kRadix + d];
    run += t4[i];
  }

  uint32_t key[kItems], val[kItems], rank[kItems];
  const uint32_t tileBase = blockIdx.x * kTile + w * kWaveTile;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t e = tileBase + j * 64 + lane;
    const bool valid = e < n;
    key[j] = valid ? keysIn[e] : 0u;
    val[j] = valid ? valsIn[e] : 0u;
  }
  __syncthreads();  // cnt zeroed, base written

  // Wave-serial ranking: slot j of the wave's 512-pair sub-tile in lane
  // order, so rank order = input order within the wave's sub-tile.
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t e = tileBase + j * 64 + lane;
    const bool valid = e < n;
    const uint32_t d = (key[j] >> shift) & (kRadix - 1u);
    const uint64_t peers = matchDigit(d, __ballot(valid));
    const uint32_t before = cnt[w][d];
    rank[j] = before + uint32_t(__popcll(peers & laneMaskLt()));
    __builtin_amdgcn_wave_barrier();
    const int leader = peers ? __builtin_ctzll(peers) : -1;
    if (valid && lane == leader) cnt[w][d] = before + uint32_t(__popcll(peers));
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // wave-exclusive offsets per digit
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const uint32_t d = threadIdx.x * kPer + i;
    uint32_t r = 0;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) {
      const uint32_t c = cnt[ww][d];
      cnt[ww][d] = r;
      r += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t e = tileBase + j * 64 + lane;
    if (e < n) {
      const uint32_t d = (key[j] >> shift) & (kRadix - 1u);
      const uint32_t dest = base[d] + cnt[w][d] + rank[j];
      if (dest < n) {  // always true for a consistent histogram; guards the writes
        keysOut[dest] = key[j];
        valsOut[dest] = val[j];
      }
    }
  }
}

}  // namespace

// Sorts c->sortKey[0] / c->sortVal[0] (n pairs); outIdx names the buffer
// pair holding the result.
int radixSort(Ctx* c, size_t n, int& outIdx) {
  outIdx = 0;
  if (n == 0) return LF_OK;
  const uint32_t nTiles = uint32_t((n + kTile - 1) / kTile);
  int rc = ensure(c, c->sortHist, size_t(nTiles) * kRadix + kRadix, "sort histogram");
  if (rc) return rc;
  uint32_t* hist = c->sortHist.p;
  uint32_t* tot = c->sortHist.p + size_t(nTiles) * kRadix;
  int src = 0;
  for (int pass = 0; pass < kPasses; ++pass) {
    const int shift = pass * kDigitBits;
    rc = launch(c, "radix_hist", k_radix_hist, dim3(nTiles), dim3(kHistThreads), 0,
                (const uint32_t*)c->sortKey[src].p, uint32_t(n), shift, hist, nTiles);
    if (rc) return rc;
    rc = launch(c, "radix_scan", k_radix_scan, dim3(kRadix / kScanDigits), dim3(kScanThreads), 0,
                hist, nTiles, tot);
    if (rc) return rc;
    rc = launch(c, "radix_scatter", k_radix_scatter, dim3(nTiles), dim3(kThreads), 0,
                (const uint32_t*)c->sortKey[src].p, (const uint32_t*)c->sortVal[src].p,
                c->sortKey[1 - src].p, c->sortVal[1 - src].p, uint32_t(n), shift,
                (const uint32_t*)hist, (const uint32_t*)tot);
    if (rc) return rc;
    src = 1 - src;
  }
  outIdx = src;
  return LF_OK;
}

}  // namespace lf
