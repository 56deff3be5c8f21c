// Neighbour-search grid on gfx950: predict + Morton key, stable radix sort
// (lf_sort.hip), payload reorder and the dense cell-start table.
//
// Replaces the host pre-pass of ocl::SphSolver::advance: advectAndCopy
// (oclsph.hpp:289-325), computeBoundAndZindex (:327-387), the sort (:733-743)
// and the grid table (:746-767), so the whole step stays on the device.
#include <algorithm>

#include "lf_internal.hpp"

namespace lf {
namespace {

constexpr int kThreads = 256;

// K1: p* and its sort key per particle (state order). The sort key is the
// scene's table base plus its Morton cell key (clamped to G in a batch of
// several scenes, so scenes never overlap; a batch of one keeps the exact key).
template <int BS>
__global__ __launch_bounds__(BS) void k_predict_key(Work W, const float4* __restrict__ X,
                                                    const float4* __restrict__ V,
                                                    const uint8_t* __restrict__ T,
                                                    const lf_well* __restrict__ wells, int nWells,
                                                    uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals) {
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    const uint32_t i = ch.begin + threadIdx.x;
    if (i >= ch.end) return;
    float px, py, pz;
    predictPstar(c, X[i], V[i], T[i], wells, nWells, px, py, pz);
    const uint32_t key = cellKey(c, px, py, pz);
    keys[i] = c.keyBase + (c.clampKeys ? min(key, c.G) : key);
    vals[i] = i;
  });
}

// K3: gather the state into sorted order; p* is recomputed from the gathered
// x, v (same arithmetic as K1, so identical bits) instead of being stored by
// K1 and gathered, and with it the scene-local cell key the pair passes use
// and P0, p* clamped to the bounds, from which the neighbour lists are built.
// v stays the pre-gravity velocity: the reference hands the device the
// velocity copied before the gravity update (oclsph.hpp:300,831).
template <int BS>
__global__ __launch_bounds__(BS) void k_reorder(
    Work W, uint32_t n, const uint32_t* __restrict__ perm, const float4* __restrict__ Xin,
    const float4* __restrict__ Vin, const uint8_t* __restrict__ Tin,
    const uint32_t* __restrict__ IDin, const lf_well* __restrict__ wells, int nWells,
    float4* __restrict__ Xout, float4* __restrict__ Vout, uint8_t* __restrict__ Tout,
    uint32_t* __restrict__ IDout, float4* __restrict__ P, uint32_t* __restrict__ localKey,
    float4* __restrict__ P0) {
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    const uint32_t s = ch.begin + threadIdx.x;
    if (s >= ch.end) return;
    const uint32_t i = min(perm[s], n - 1u);  // perm is a permutation; the clamp only bounds reads
    const float4 x = Xin[i];
    const float4 v = Vin[i];
    const uint8_t t = Tin[i];
    float px, py, pz;
    predictPstar(c, x, v, t, wells, nWells, px, py, pz);
    Xout[s] = x;
    Vout[s] = v;
    Tout[s] = t;
    IDout[s] = IDin[i];
    P[s] = make_float4(px, py, pz, x.w);
    localKey[s] = cellKey(c, px, py, pz);
    P0[s] = make_float4(clampToBounds(c, 0, px), clampToBounds(c, 1, py), clampToBounds(c, 2, pz),
                        x.w);
  });
}

// Lower bound of key k in sorted keys[0, n).
__device__ inline uint32_t lowerBound(const uint32_t* __restrict__ keys, uint32_t n, uint32_t k) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (keys[mid] < k)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

constexpr uint32_t kCells = 1024;            // cells per workgroup
constexpr uint32_t kRounds = kCells / kThreads;
constexpr uint32_t kMarkUnroll = 4;          // keys in flight per thread while marking

// K4: table[c] = first sorted index whose key >= c, for c in [0, G) (G: the
// concatenated tables of all scenes): the monotone scan of oclsph.hpp:758-765
// as a per-tile lower bound. Each workgroup owns kCells cells: the run starts
// of its key range are marked in LDS, then a suffix minimum fills each empty
// cell with the next start. The suffix minimum runs over rounds of 256 cells,
// last round first, each a wave-shuffle suffix minimum plus the minima of the
// later waves and rounds, so LDS reads and table stores are consecutive.
__global__ __launch_bounds__(kThreads) void k_cell_table(const uint32_t* __restrict__ keys,
                                                         uint32_t n, uint32_t G,
                                                         uint32_t* __restrict__ table,
                                                         uint32_t* __restrict__ cellDisp) {
  __shared__ uint32_t local[kCells];
  __shared__ uint32_t range[2];
  __shared__ uint32_t waveMin[kThreads / 64];
  const uint32_t c0 = blockIdx.x * kCells;
  const uint32_t c1 = min(c0 + kCells, G);
  if (threadIdx.x < 2) range[threadIdx.x] = lowerBound(keys, n, threadIdx.x == 0 ? c0 : c1);
  __syncthreads();
  const uint32_t s0 = range[0], s1 = range[1];
  for (uint32_t i = threadIdx.x; i < kCells; i += kThreads) local[i] = s1;
  __syncthreads();
  // run starts; a dense tile holds ~8 particles per cell, so each thread
  // walks its strided share with kMarkUnroll key pairs in flight
  for (uint32_t base = s0; base < s1; base += kThreads * kMarkUnroll) {
    uint32_t k[kMarkUnroll], kp[kMarkUnroll];
#pragma unroll
    for (uint32_t u = 0; u < kMarkUnroll; ++u) {
      const uint32_t s = min(base + u * kThreads + threadIdx.x, s1 - 1u);
      k[u] = keys[s];
      kp[u] = s > 0 ? keys[s - 1] : ~k[u];
    }
#pragma unroll
    for (uint32_t u = 0; u < kMarkUnroll; ++u) {
      const uint32_t s = base + u * kThreads + threadIdx.x;
      if (s < s1 && kp[u] != k[u] && k[u] - c0 < kCells) local[k[u] - c0] = s;
    }
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  uint32_t carry = s1;  // minimum over the cells of the rounds already done
  for (int r = int(kRounds) - 1; r >= 0; --r) {
    const uint32_t i = uint32_t(r) * kThreads + threadIdx.x;
    uint32_t v = local[i];
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
      const uint32_t o = __shfl_down(v, off, 64);
      if (lane + off < 64) v = min(v, o);
    }
    if (lane == 0) waveMin[w] = v;  // lane 0 holds its wave's minimum
    __syncthreads();
    uint32_t later = carry;
    for (uint32_t x = w + 1; x < kThreads / 64; ++x) later = min(later, waveMin[x]);
    const uint32_t cell = c0 + i;
    if (cell < c1) {
      table[cell] = min(v, later);
      cellDisp[cell] = 0u;  // per-cell displacement of the neighbour lists
    }
    for (uint32_t x = 0; x < kThreads / 64; ++x) carry = min(carry, waveMin[x]);
    __syncthreads();
  }
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
  const uint32_t G = c->tableTotal;
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
  if ((rc = ensure(c, c->localKey, n, "cell keys"))) return rc;
  if ((rc = ensure(c, c->nbr, nbrCapacity(n), "neighbour lists"))) return rc;
  if ((rc = ensure(c, c->nbrCount, n, "neighbour counts"))) return rc;
  if ((rc = ensure(c, c->P0, n, "list positions"))) return rc;
  if ((rc = ensure(c, c->maxDisp, 4, "displacement"))) return rc;
  if ((rc = ensure(c, c->rho, n, "density"))) return rc;
  if ((rc = ensure(c, c->diffused, n, "diffused"))) return rc;
  if ((rc = ensure(c, c->table, size_t(G), "cell table"))) return rc;
  if ((rc = ensure(c, c->tableAlt, size_t(G), "cell table"))) return rc;
  if ((rc = ensure(c, c->cellDisp, size_t(G), "cell displacement"))) return rc;
  if (n == 0) return LF_OK;
  const Work W = workOf(c, -1);
  rc = LF_DISPATCH_BS(c, "predict_key", k_predict_key, W, (const float4*)c->X[c->cur].p,
                      (const float4*)c->V[c->cur].p, (const uint8_t*)c->T[c->cur].p,
                      (const lf_well*)c->wells.p, c->nWells, c->sortKey[0].p, c->sortVal[0].p);
  if (rc) return rc;
  int idx = 0;
  if ((rc = radixSort(c, n, idx))) return rc;
  rc = LF_DISPATCH_BS(c, "reorder", k_reorder, W, uint32_t(n), (const uint32_t*)c->sortVal[idx].p,
                      (const float4*)c->X[c->cur].p, (const float4*)c->V[c->cur].p,
                      (const uint8_t*)c->T[c->cur].p, (const uint32_t*)c->ID[c->cur].p,
                      (const lf_well*)c->wells.p, c->nWells, c->X[nxt].p, c->V[nxt].p,
                      c->T[nxt].p, c->ID[nxt].p, c->P.p, c->localKey.p, c->P0.p);
  if (rc) return rc;
  rc = launch(c, "cell_table", k_cell_table, dim3(blocksFor(G, kCells)), dim3(kThreads), 0,
              (const uint32_t*)c->sortKey[idx].p, uint32_t(n), G, c->table.p, c->cellDisp.p);
  if (rc) return rc;
  notePositionsWritten(c);  // P (reorder) and the table
  c->keysIdx = idx;
  c->gridG = c->consts.G;
  c->gridTotal = G;
  c->cur = nxt;
  c->gridValid = true;
  return LF_OK;
}

}  // namespace lf
