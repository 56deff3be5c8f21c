// Stable LSD radix sort of (Morton key, particle index) pairs on gfx950.
//
// Replaces the host std::sort of oclsph.hpp:733-743. The reference sort is
// unstable; this one is stable, so ties keep the pre-sort order (the CPU
// restatement uses std::stable_sort for the same order).
//
// Keys are 30-bit Morton codes (10 bits per axis, curves.h:56-72): three
// passes of 10-bit digits. Each pass is three kernels over tiles of 2048
// pairs (256 threads x 8):
//   hist    per-tile digit counts (wave-aggregated LDS counters), digit-major
//   scan    per-digit exclusive scan over tiles (one workgroup per digit)
//   scatter digit bases + stable in-tile rank (per-wave ballot match, LDS
//           running counters in (wave, slot, lane) = input order) -> output
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

__global__ __launch_bounds__(kThreads) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                         uint32_t n, int shift,
                                                         uint32_t* __restrict__ hist,
                                                         uint32_t nTiles) {
  __shared__ uint32_t h[kRadix];
  for (uint32_t i = threadIdx.x; i < kRadix; i += kThreads) h[i] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kTile;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t e = base + j * kThreads + threadIdx.x;
    const bool valid = e < n;
    const uint32_t d = valid ? (keys[e] >> shift) & (kRadix - 1u) : 0u;
    const uint64_t peers = matchDigit(d, __ballot(valid));
    const int leader = peers ? __builtin_ctzll(peers) : -1;
    if (valid && __lane_id() == leader) atomicAdd(&h[d], uint32_t(__popcll(peers)));
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kRadix; i += kThreads) hist[i * nTiles + blockIdx.x] = h[i];
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

// One workgroup per digit: exclusive scan of hist[d][0..nTiles) in place,
// digit total into tot[d].
__global__ __launch_bounds__(kThreads) void k_radix_scan(uint32_t* __restrict__ hist,
                                                         uint32_t nTiles,
                                                         uint32_t* __restrict__ tot) {
  __shared__ uint32_t sm[kWaves];
  uint32_t* row = hist + size_t(blockIdx.x) * nTiles;
  const uint32_t per = (nTiles + kThreads - 1) / kThreads;
  const uint32_t lo = threadIdx.x * per;
  const uint32_t hi = min(lo + per, nTiles);
  uint32_t s = 0;
  for (uint32_t i = lo; i < hi; ++i) s += row[i];
  uint32_t total;
  uint32_t run = blockExclusiveScan(s, sm, &total);
  for (uint32_t i = lo; i < hi; ++i) {
    const uint32_t t = row[i];
    row[i] = run;
    run += t;
  }
  if (threadIdx.x == 0) tot[blockIdx.x] = total;
}

__global__ __launch_bounds__(kThreads) void k_radix_scatter(
    const uint32_t* __restrict__ keysIn, const uint32_t* __restrict__ valsIn,
    uint32_t* __restrict__ keysOut, uint32_t* __restrict__ valsOut, uint32_t n, int shift,
    const uint32_t* __restrict__ hist, const uint32_t* __restrict__ tot, uint32_t nTiles) {
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
    base[d] = run + hist[size_t(d) * nTiles + blockIdx.x];
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
    rc = launch(c, "radix_hist", k_radix_hist, dim3(nTiles), dim3(kThreads), 0,
                (const uint32_t*)c->sortKey[src].p, uint32_t(n), shift, hist, nTiles);
    if (rc) return rc;
    rc = launch(c, "radix_scan", k_radix_scan, dim3(kRadix), dim3(kThreads), 0, hist, nTiles,
                tot);
    if (rc) return rc;
    rc = launch(c, "radix_scatter", k_radix_scatter, dim3(nTiles), dim3(kThreads), 0,
                (const uint32_t*)c->sortKey[src].p, (const uint32_t*)c->sortVal[src].p,
                c->sortKey[1 - src].p, c->sortVal[1 - src].p, uint32_t(n), shift,
                (const uint32_t*)hist, (const uint32_t*)tot, nTiles);
    if (rc) return rc;
    src = 1 - src;
  }
  outIdx = src;
  return LF_OK;
}

}  // namespace lf
