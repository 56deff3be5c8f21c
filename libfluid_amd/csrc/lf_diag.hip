// FP64 per-scene reductions on gfx950: conserved quantities (fluid mass and
// momentum) and the density error |rho/rho0 - 1| of the last lambda pass.
// No reference counterpart (libfluid reports none).
//
// Two kernels, both in a fixed order so the sums are reproducible run to run:
// k_chunk_sums reduces each chunk (<= one workgroup of particles of one
// scene) in FP64 with a fixed tree; k_scene_sums adds a scene's chunk partials
// with one workgroup per scene (strided partials, then a fixed tree).
// densityCheck also drives the convergence check: a scene whose mean density
// error falls below its density_tol is flagged, and the remaining lambda /
// delta passes of the step skip its chunks (all on the device, so a captured
// graph keeps working).
#include <type_traits>

#include "lf_internal.hpp"

namespace lf {
namespace {

constexpr int kSceneThreads = 256;

__device__ inline double waveSum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double waveMax(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Per chunk: [0] fluid count, [1] sum m, [2..4] sum m v, [5] sum |e|,
// [6] max |e| with e = rho/rho0 - 1 (rho of the last lambda pass).
template <int BS>
__global__ __launch_bounds__(BS) void k_chunk_sums(Work W, const uint8_t* __restrict__ T,
                                                   const float4* __restrict__ X,
                                                   const float4* __restrict__ V,
                                                   const float* __restrict__ rho,
                                                   double* __restrict__ out) {
  __shared__ double part[BS / 64][kSumWidth];
  forEachChunk(W, [&](const PbfConsts& c, const Chunk& ch) {
    const uint32_t a = ch.begin + threadIdx.x;
    double v[kSumWidth] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (a < ch.end && T[a] != LF_OBSTACLE) {
      const float4 x = X[a];
      const float4 u = V[a];
      const double m = double(x.w);
      v[0] = 1.0;
      v[1] = m;
      v[2] = m * double(u.x);
      v[3] = m * double(u.y);
      v[4] = m * double(u.z);
      const double e = fabs(double(rho[a]) / double(c.rho0) - 1.0);
      v[5] = e;
      v[6] = e;
    }
#pragma unroll
    for (int k = 0; k < kSumWidth; ++k) v[k] = (k == 6) ? waveMax(v[k]) : waveSum(v[k]);
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0)
#pragma unroll
      for (int k = 0; k < kSumWidth; ++k) part[w][k] = v[k];
    __syncthreads();
    if (threadIdx.x < kSumWidth) {
      const int k = threadIdx.x;
      double r = part[0][k];
      for (int i = 1; i < BS / 64; ++i) r = (k == 6) ? fmax(r, part[i][k]) : r + part[i][k];
      out[size_t(ch.index) * kSumWidth + k] = r;
    }
    __syncthreads();
  });
}

// One workgroup per scene: adds the scene's chunk partials into
// stats[scene][slot]. With pass >= 0 (after lambda pass `pass`), also updates
// the convergence flag and the pass count of scenes with density_tol > 0.
__global__ __launch_bounds__(kSceneThreads) void k_scene_sums(
    const double* __restrict__ chunkSums, const uint32_t* __restrict__ chunkOff,
    const PbfConsts* __restrict__ consts, int pass, double* __restrict__ stats,
    uint32_t* __restrict__ done, uint32_t* __restrict__ passes) {
  __shared__ double part[kSceneThreads / 64][kSumWidth];
  const uint32_t s = blockIdx.x;
  if (pass >= 0) {
    // a scene that converged before this pass did not run it
    if (pass > 0 && done[s]) return;
  }
  const uint32_t c0 = chunkOff[s], c1 = chunkOff[s + 1];
  double v[kSumWidth] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (uint32_t i = c0 + threadIdx.x; i < c1; i += kSceneThreads) {
    const double* p = chunkSums + size_t(i) * kSumWidth;
#pragma unroll
    for (int k = 0; k < kSumWidth; ++k) v[k] = (k == 6) ? fmax(v[k], p[k]) : v[k] + p[k];
  }
#pragma unroll
  for (int k = 0; k < kSumWidth; ++k) v[k] = (k == 6) ? waveMax(v[k]) : waveSum(v[k]);
  const int w = threadIdx.x >> 6;
  if (__lane_id() == 0)
#pragma unroll
    for (int k = 0; k < kSumWidth; ++k) part[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x != 0) return;
  double r[kSumWidth];
#pragma unroll
  for (int k = 0; k < kSumWidth; ++k) {
    r[k] = part[0][k];
    for (int i = 1; i < kSceneThreads / 64; ++i) r[k] = (k == 6) ? fmax(r[k], part[i][k]) : r[k] + part[i][k];
  }
  if (pass < 0) {  // state slot
    double* o = stats + (size_t(s) * kStatSlots + kPassSlots) * kSumWidth;
#pragma unroll
    for (int k = 0; k < kSumWidth; ++k) o[k] = r[k];
    return;
  }
  double* row = stats + size_t(s) * kStatSlots * kSumWidth;
  if (pass == 0) {
    for (int i = 0; i < kPassSlots; ++i) row[size_t(i) * kSumWidth + 5] = 0.0;
    done[s] = 0u;
  }
  const double mean = r[0] > 0.0 ? r[5] / r[0] : 0.0;
  if (pass < kPassSlots) row[size_t(pass) * kSumWidth + 5] = mean;
  passes[s] = uint32_t(pass + 1);
  const float tol = consts[s].densityTol;
  if (tol > 0.0f && mean < double(tol)) done[s] = 1u;
}

template <typename F> int withBS(const Ctx* c, F&& f) {
  switch (c->chunkThreads) {
    case 128: return f(std::integral_constant<int, 128>{});
    case 512: return f(std::integral_constant<int, 512>{});
    default: return f(std::integral_constant<int, 256>{});
  }
}

int reduce(Ctx* c, int pass) {
  if (c->n == 0 || c->nChunks == 0) return LF_OK;
  const Work W = workOf(c, -1);
  int rc = withBS(c, [&](auto bs) {
    return launch(c, "chunk_sums", k_chunk_sums<decltype(bs)::value>, gridOf(c, W),
                  dim3(decltype(bs)::value), 0, W, (const uint8_t*)c->T[c->cur].p,
                  (const float4*)c->X[c->cur].p, (const float4*)c->V[c->cur].p,
                  (const float*)c->rho.p, c->chunkSums.p);
  });
  if (rc) return rc;
  const uint32_t S = uint32_t(c->chunkOffHost.size()) - 1u;
  return launch(c, "scene_sums", k_scene_sums, dim3(S), dim3(kSceneThreads), 0,
                (const double*)c->chunkSums.p, (const uint32_t*)c->chunkOff.p,
                (const PbfConsts*)c->dConsts.p, pass, c->sceneStats.p, c->sceneDone.p,
                c->scenePasses.p);
}

}  // namespace

int densityCheck(Ctx* c, int pass) { return reduce(c, pass); }
int diagState(Ctx* c) { return reduce(c, -1); }

}  // namespace lf
