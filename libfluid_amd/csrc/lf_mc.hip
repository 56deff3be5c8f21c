// Marching cubes on gfx950: field sampling, case classification + triangle
// count, a device-side scan of the counts, and triangle emission at the
// scanned offsets (cube-index order, so the output is deterministic; the
// reference appends with atomic_inc, oclsph_kernel.h:713).
//
// Lattice layout: sample (x,y,z) at index (x*Sy + y)*Sz + z (index3d,
// curves.h:17-20); cube (x,y,z) at (x*My + y)*Mz + z with M = S - 1 (to3d,
// curves.h:22-27). Corner order: CUBE_OFFSETS of oclsph_kernel.h:533-542.
#include <mutex>

#include "fluid/mctable.hpp"
#include "lf_internal.hpp"

namespace lf {
namespace {

constexpr int kThreads = 256;
constexpr int kCubesPerThread = 4;
constexpr uint32_t kCubesPerBlock = kThreads * kCubesPerThread;

struct McTables {
  uint16_t edge[256];
  uint8_t ntri[256];
  uint8_t tri[256][16];  // edge indices, 255 terminated
};
constexpr McTables makeTables() {
  McTables t{};
  for (int ci = 0; ci < 256; ++ci) {
    t.edge[ci] = uint16_t(mctable::EdgeTable[ci]);
    t.ntri[ci] = uint8_t(mctable::EdgeTable[ci] == 0 ? 0 : mctable::NumVertsTable[ci] / 3);
    for (int i = 0; i < 16; ++i)
      t.tri[ci][i] = uint8_t(mctable::TriTable[ci][i] < 0 ? 255 : mctable::TriTable[ci][i]);
  }
  return t;
}
__constant__ McTables cTab = makeTables();

__constant__ int cCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
__constant__ int cEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                 {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

struct Lattice {
  uint32_t S[3];
  float origin[3];  // solver units
  float step[3];
};

// ---- field sampling -----------------------------------------------------------

// SPH density field rho(x) = sum over fluid b of m_b poly6 (h^2 - d^2)^3 for
// d^2 = |x - p*_b|^2 <= h^2 (W6 written on d^2, no square root); P.w is the
// fluid mass (0 for obstacles) after finalise, staged as m_b poly6.
// Candidates: the 27 cells around the sample's own cell, reference order
// (z rows outer, then y rows, each row the x-run of three cells).
//
// One workgroup per tile of samples: the particles of the tile's halo cells
// (its cell range +-1 per axis) are staged once in LDS with a local cell
// table, so the candidate loops read LDS; a tile whose halo holds no particle
// writes zeros and exits. Tiles whose halo does not fit (coarse lattices,
// crowded cells) read global memory instead, same order.
//
// Each lane owns a column of kNZ samples along z, and each wave a kWX x kWY
// block of columns over the same z range, so the samples' z cells are the
// same on every lane of a wave. The lane walks the union of its samples' z
// rows once: per candidate one LDS read and the x and y part of the distance,
// u = fma(-dy, dy, fma(-dx, dx, h^2)), shared by the column, then per sample
// whose 27 cells contain the row t = max(fma(-dz, dz, u), 0) and
// rho = fma(w_b t, t t, rho). Which samples a row belongs to is the same on
// every lane, so the loop for each row is chosen once (a uniform branch).
// Each sample still sees exactly its 27 cells in the reference order.
// Outside h, and for a NaN or infinite distance, t is 0 before it is used, so
// the term is exactly 0, as the test d2 <= h^2 would make it.
#ifndef LF_MC_UNROLL
#define LF_MC_UNROLL 4
#endif
constexpr uint32_t kUnroll = LF_MC_UNROLL;  // staged particles in flight per lane
constexpr int kHaloCells = 512;

// The x and y part of a candidate's distance term, shared by a column.
__device__ inline float columnPart(float hh, float4 q, float px, float py) {
  const float dx = px - q.x, dy = py - q.y;
  return __builtin_fmaf(-dy, dy, __builtin_fmaf(-dx, dx, hh));
}
// One sample's term from the shared part.
__device__ inline void addTerm(float& rho, float u, float4 q, float pz) {
  const float dz = pz - q.z;
  const float t = fmaxf(__builtin_fmaf(-dz, dz, u), 0.0f);
  rho = __builtin_fmaf(q.w * t, t * t, rho);
}

// The terms of staged candidates [b, e) for the column's samples KLO..KHI.
template <int KLO, int KHI, int kNZ>
__device__ inline void rowTerms(const float4* __restrict__ staged, uint32_t b, uint32_t e, float hh,
                                float px, float py, const float (&pz)[kNZ], float (&rho)[kNZ]) {
  for (; b + kUnroll <= e; b += kUnroll) {
    float4 q[kUnroll];
#pragma unroll
    for (uint32_t v = 0; v < kUnroll; ++v) q[v] = staged[b + v];
#pragma unroll
    for (uint32_t v = 0; v < kUnroll; ++v) {
      const float u = columnPart(hh, q[v], px, py);
#pragma unroll
      for (int k = KLO; k <= KHI; ++k) addTerm(rho[k], u, q[v], pz[k]);
    }
  }
  for (; b < e; ++b) {
    const float4 q = staged[b];
    const float u = columnPart(hh, q, px, py);
#pragma unroll
    for (int k = KLO; k <= KHI; ++k) addTerm(rho[k], u, q, pz[k]);
  }
}

// Dispatch on the (uniform) range of samples a row belongs to.
template <int kNZ, int KLO = 0, int KHI = 0>
__device__ inline void rowTermsFor(int lo, int hi, const float4* __restrict__ staged, uint32_t b,
                                   uint32_t e, float hh, float px, float py,
                                   const float (&pz)[kNZ], float (&rho)[kNZ]) {
  if constexpr (KLO < kNZ) {
    if constexpr (KHI < kNZ) {
      if (lo == KLO && hi == KHI) {
        rowTerms<KLO, KHI, kNZ>(staged, b, e, hh, px, py, pz, rho);
        return;
      }
      rowTermsFor<kNZ, KLO, KHI + 1>(lo, hi, staged, b, e, hh, px, py, pz, rho);
    } else {
      rowTermsFor<kNZ, KLO + 1, KLO + 1>(lo, hi, staged, b, e, hh, px, py, pz, rho);
    }
  }
}

__device__ inline void cellCoords(const PbfConsts& c, float px, float py, float pz, uint32_t& x,
                                  uint32_t& y, uint32_t& z) {
  x = cellCoord((px - c.minExt[0]) / c.h);
  y = cellCoord((py - c.minExt[1]) / c.h);
  z = cellCoord((pz - c.minExt[2]) / c.h);
}

// Tile: kWX x kWY columns (one wave) x kNZ * kWaves samples along z (the
// waves stacked in z); kCap staged particles.
template <int kWX, int kWY, int kNZ, int kWaves, int kCap>
__global__ __launch_bounds__(64 * kWaves) void k_mc_density(Lattice L, PbfConsts c,
                                                             const uint32_t* __restrict__ table,
                                                             const float4* __restrict__ P,
                                                             float* __restrict__ field,
                                                             uint32_t tilesY, uint32_t tilesZ,
                                                             const uint32_t* __restrict__ health,
                                                             uint8_t* __restrict__ tileEmpty) {
  static_assert(kWX * kWY == 64, "a wave is a kWX x kWY block of columns");
  constexpr uint32_t kT = 64 * kWaves;
  constexpr uint32_t kTX = kWX, kTY = kWY, kTZ = kNZ * kWaves;
  constexpr uint32_t kRounds = (kHaloCells + kT - 1) / kT;  // halo cells per thread
  if (health[0] | health[1]) return;  // a diverged state: no field
  __shared__ float4 staged[kCap];  // (x, y, z, m poly6): one LDS read per candidate
  __shared__ uint32_t cellStart[kHaloCells + 1];
  __shared__ uint32_t scanTmp[kWaves];
  __shared__ int mode;  // 0 LDS, 1 global
  const uint32_t tz = blockIdx.x % tilesZ;
  const uint32_t ty = (blockIdx.x / tilesZ) % tilesY;
  const uint32_t tx = blockIdx.x / (tilesZ * tilesY);
  const uint32_t x0 = tx * kTX, y0 = ty * kTY, z0 = tz * kTZ;
  const uint32_t x1 = min(x0 + kTX, L.S[0]) - 1, y1 = min(y0 + kTY, L.S[1]) - 1,
                 z1 = min(z0 + kTZ, L.S[2]) - 1;
  // the lane's column
  const uint32_t w = threadIdx.x >> 6, li = threadIdx.x & 63u;
  const uint32_t x = x0 + li / kWY, y = y0 + li % kWY, zb = z0 + w * kNZ;
  const bool inside = x < L.S[0] && y < L.S[1] && zb < L.S[2];
  const uint64_t col = (uint64_t(x) * L.S[1] + y) * L.S[2] + zb;
  auto writeColumn = [&](const float (&v)[kNZ]) {
#pragma unroll
    for (int k = 0; k < kNZ; ++k)
      if (zb + k < L.S[2]) field[col + k] = v[k];
  };
  // cell range of the tile (cellCoord is monotone in the sample position)
  uint32_t lo[3], hi[3];
  cellCoords(c, __builtin_fmaf(float(x0), L.step[0], L.origin[0]),
             __builtin_fmaf(float(y0), L.step[1], L.origin[1]),
             __builtin_fmaf(float(z0), L.step[2], L.origin[2]), lo[0], lo[1], lo[2]);
  cellCoords(c, __builtin_fmaf(float(x1), L.step[0], L.origin[0]),
             __builtin_fmaf(float(y1), L.step[1], L.origin[1]),
             __builtin_fmaf(float(z1), L.step[2], L.origin[2]), hi[0], hi[1], hi[2]);
  bool fits = true;
  uint32_t hx = 0, hy = 0, hz = 0;
  for (int k = 0; k < 3; ++k) fits = fits && lo[k] >= 1 && hi[k] <= 1022;
  if (fits) {
    hx = hi[0] - lo[0] + 3;
    hy = hi[1] - lo[1] + 3;
    hz = hi[2] - lo[2] + 3;
    fits = hx * hy * hz <= uint32_t(kHaloCells);
  }
  const uint32_t nh = fits ? hx * hy * hz : 0;
  if (fits) {
    // Quick emptiness test: an aligned 4x4x4 block of cells is one run of 64
    // Morton codes, so table[end] - table[start] counts its particles (codes
    // up to G - 1, the cells that hold any). A tile whose halo lies in blocks
    // holding none writes zeros without staging anything. (More blocks than
    // threads: no test.)
    const uint32_t b0[3] = {(lo[0] - 1) >> 2, (lo[1] - 1) >> 2, (lo[2] - 1) >> 2};
    const uint32_t nb[3] = {((hi[0] + 1) >> 2) - b0[0] + 1, ((hi[1] + 1) >> 2) - b0[1] + 1,
                            ((hi[2] + 1) >> 2) - b0[2] + 1};
    const uint32_t nbl = nb[0] * nb[1] * nb[2];
    bool any = nbl > kT;
    if (threadIdx.x < nbl && !any) {
      const uint32_t i = threadIdx.x;
      const uint32_t base = mortonEncode(4u * (b0[0] + i % nb[0]), 4u * (b0[1] + (i / nb[0]) % nb[1]),
                                         4u * (b0[2] + i / (nb[0] * nb[1])));
      const uint32_t last = c.G - 1u;
      any = table[min(base + 64u, last)] != table[min(base, last)];
    }
    if (!__syncthreads_or(any)) {
      if (threadIdx.x == 0) tileEmpty[blockIdx.x] = 1u;
      if (inside) {
        float z[kNZ];
#pragma unroll
        for (int k = 0; k < kNZ; ++k) z[k] = 0.0f;
        writeColumn(z);
      }
      return;
    }
  }
  // count the halo cells' particles (per thread: cells tid + r * kT)
  uint32_t myCount[kRounds];
#pragma unroll
  for (uint32_t r = 0; r < kRounds; ++r) {
    myCount[r] = 0;
    const uint32_t h = threadIdx.x + r * kT;
    if (h < nh) {
      const uint32_t ix = h % hx, iy = (h / hx) % hy, iz = h / (hx * hy);
      const uint32_t code = mortonEncode(lo[0] - 1 + ix, lo[1] - 1 + iy, lo[2] - 1 + iz);
      if (code + 1u < c.G) myCount[r] = table[code + 1u] - table[code];
    }
  }
  // exclusive scan over cells in order h = 0..nh-1, one round of kT cells at
  // a time
  uint32_t totalParticles = 0;
  {
    const uint32_t lane = __lane_id();
#pragma unroll
    for (uint32_t r = 0; r < kRounds; ++r) {
      const uint32_t v = myCount[r];
      uint32_t xs = v;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t yv = __shfl_up(xs, o, 64);
        if (lane >= uint32_t(o)) xs += yv;
      }
      if (lane == 63u) scanTmp[w] = xs;
      __syncthreads();
      uint32_t wb = 0, tot = 0;
#pragma unroll
      for (uint32_t i = 0; i < kWaves; ++i) {
        if (i < w) wb += scanTmp[i];
        tot += scanTmp[i];
      }
      const uint32_t h = threadIdx.x + r * kT;
      if (h < nh) cellStart[h] = totalParticles + wb + xs - v;
      totalParticles += tot;
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      cellStart[nh] = totalParticles;
      mode = (!fits || totalParticles > uint32_t(kCap)) ? 1 : 0;
    }
  }
  __syncthreads();
  const bool useLds = mode == 0;
  if (useLds) {
#pragma unroll
    for (uint32_t r = 0; r < kRounds; ++r) {
      const uint32_t h = threadIdx.x + r * kT;
      if (h < nh && myCount[r] > 0) {
        const uint32_t ix = h % hx, iy = (h / hx) % hy, iz = h / (hx * hy);
        const uint32_t code = mortonEncode(lo[0] - 1 + ix, lo[1] - 1 + iy, lo[2] - 1 + iz);
        const uint32_t s0 = table[code], dst = cellStart[h], cnt = myCount[r];
        // four loads in flight per round (the last round's extra lanes load
        // the cell's last particle again and store nothing)
        for (uint32_t j = 0; j < cnt; j += 4) {
          float4 q[4];
#pragma unroll
          for (uint32_t u = 0; u < 4; ++u) q[u] = P[s0 + min(j + u, cnt - 1u)];
#pragma unroll
          for (uint32_t u = 0; u < 4; ++u)
            if (j + u < cnt) staged[dst + j + u] = make_float4(q[u].x, q[u].y, q[u].z, q[u].w * c.poly6);
        }
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) tileEmpty[blockIdx.x] = 0u;
  if (!inside) return;
  // samples past the lattice's end repeat its last one (computed, not written),
  // so every sample's cells lie in the tile's range
  float rho[kNZ], pz[kNZ];
#pragma unroll
  for (int k = 0; k < kNZ; ++k) {
    rho[k] = 0.0f;
    pz[k] = __builtin_fmaf(float(min(zb + uint32_t(k), L.S[2] - 1u)), L.step[2], L.origin[2]);
  }
  if (useLds && totalParticles == 0) {
    writeColumn(rho);
    return;
  }
  const float px = __builtin_fmaf(float(x), L.step[0], L.origin[0]);
  const float py = __builtin_fmaf(float(y), L.step[1], L.origin[1]);
  if (useLds) {
    uint32_t cx, cy, cz0;
    cellCoords(c, px, py, pz[0], cx, cy, cz0);
    // the z cells of the column's samples: the same on every lane of the
    // wave (they depend on z only); halo z index of each sample's dz = -1 row
    int zl[kNZ];
#pragma unroll
    for (int k = 0; k < kNZ; ++k)
      zl[k] = __builtin_amdgcn_readfirstlane(int(cellCoord((pz[k] - c.minExt[2]) / c.h) - lo[2]));
    const uint32_t bx = cx - lo[0], by = cy - lo[1];  // halo index of dx = -1, dy = -1
    for (int iz = zl[0]; iz <= zl[kNZ - 1] + 2; ++iz) {
      // samples whose rows include halo z row iz: zl[k] <= iz <= zl[k] + 2
      int klo = kNZ, khi = -1;
#pragma unroll
      for (int k = 0; k < kNZ; ++k)
        if (zl[k] <= iz && iz <= zl[k] + 2) {
          klo = min(klo, k);
          khi = max(khi, k);
        }
      for (uint32_t dy = 0; dy < 3; ++dy) {
        // the three x cells of a row are consecutive halo cells, so their
        // particles are one contiguous run of the staging, in visit order
        const uint32_t row = (uint32_t(iz) * hy + (by + dy)) * hx + bx;
        rowTermsFor<kNZ>(klo, khi, staged, cellStart[row], cellStart[row + 3], c.hh, px, py, pz,
                         rho);
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < kNZ; ++k)
      forEachNeighbourCell(cellKey(c, px, py, pz[k]), table, c.G, [&](uint32_t s, uint32_t e) {
        for (uint32_t b = s; b < e; ++b) {
          const float4 pb = P[b];
          const float4 q = make_float4(pb.x, pb.y, pb.z, pb.w * c.poly6);
          addTerm(rho[k], columnPart(c.hh, q, px, py), q, pz[k]);
        }
      });
  }
  writeColumn(rho);
}

// mc_lattice (oclsph_kernel.h:421-530): metaball field sum of size/|l|^infl,
// normal, mean colour; 27 cells clamped to the extent. Undefined reference
// cases are written as zeros (the skipped last z-curve cell, colour/N at N=0).
__global__ __launch_bounds__(kThreads) void k_mc_reference(
    uint64_t total, Lattice L, PbfConsts c, float res, float psize, float infl,
    const uint32_t* __restrict__ table, const uint8_t* __restrict__ T,
    const float4* __restrict__ X, const float4* __restrict__ V, float* __restrict__ field,
    float4* __restrict__ normalOut, uint32_t* __restrict__ colourOut,
    const uint32_t* __restrict__ health) {
  const uint64_t idx = uint64_t(blockIdx.x) * kThreads + threadIdx.x;
  if (idx >= total || (health[0] | health[1])) return;  // a diverged state: no field
  const uint32_t z = uint32_t(idx % L.S[2]);
  const uint64_t xy = idx / L.S[2];
  const uint32_t y = uint32_t(xy % L.S[1]);
  const uint32_t x = uint32_t(xy / L.S[1]);
  const float step = c.h / res;
  const float ax = __builtin_fmaf(float(x), step, c.minExt[0]) * c.scale;
  const float ay = __builtin_fmaf(float(y), step, c.minExt[1]) * c.scale;
  const float az = __builtin_fmaf(float(z), step, c.minExt[2]) * c.scale;
  const uint32_t cx = cellCoord(float(x) / res), cy = cellCoord(float(y) / res),
                 cz = cellCoord(float(z) / res);
  field[idx] = 0.0f;
  normalOut[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  colourOut[idx] = 0u;
  if (cx == c.extent[0] && cy == c.extent[1] && cz == c.extent[2]) return;
  const float thr = c.h * c.scale * 1.0f;
  const float k = (-infl) * psize;
  auto clampc = [](int v, uint32_t e) { return uint32_t(min(max(v, 0), int(e) - 1)); };
  const uint32_t xs[3] = {clampc(int(cx) - 1, c.extent[0]), cx, clampc(int(cx) + 1, c.extent[0])};
  const uint32_t ys[3] = {clampc(int(cy) - 1, c.extent[1]), cy, clampc(int(cy) + 1, c.extent[1])};
  const uint32_t zs[3] = {clampc(int(cz) - 1, c.extent[2]), cz, clampc(int(cz) + 1, c.extent[2])};
  float val = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f, cc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  uint32_t N = 0;
  for (int iz = 0; iz < 3; ++iz)
    for (int iy = 0; iy < 3; ++iy)
      for (int ix = 0; ix < 3; ++ix) {
        const uint32_t code = mortonEncode(xs[ix], ys[iy], zs[iz]);
        if (code + 1u >= c.G) continue;
        const uint32_t e = table[code + 1u];
        for (uint32_t b = table[code]; b < e; ++b) {
          if (T[b] == LF_OBSTACLE) continue;
          const float4 xb = X[b];
          const float lx = xb.x - ax, ly = xb.y - ay, lz = xb.z - az;
          const float len = __builtin_sqrtf(dist2(lx, ly, lz));
          if (!(len < thr)) continue;
          const float den = (infl == 0.5f) ? __builtin_sqrtf(len) : powf(len, infl);
          nx = __builtin_fmaf(k, lx / den, nx);
          ny = __builtin_fmaf(k, ly / den, ny);
          nz = __builtin_fmaf(k, lz / den, nz);
          val += psize / den;
          const uint32_t col = __float_as_uint(V[b].w);
          cc[0] += float((col >> 24) & 0xFFu);
          cc[1] += float((col >> 16) & 0xFFu);
          cc[2] += float((col >> 8) & 0xFFu);
          cc[3] += float(col & 0xFFu);
          ++N;
        }
      }
  const float nn = dist2(nx, ny, nz);
  if (nn > 0.0f) {
    const float inv = 1.0f / __builtin_sqrtf(nn);
    nx *= inv;
    ny *= inv;
    nz *= inv;
  }
  field[idx] = val;
  normalOut[idx] = make_float4(nx, ny, nz, 0.0f);
  uint32_t out = 0;
  if (N > 0) {
    for (int q = 0; q < 4; ++q) {
      const float m = cc[q] / float(N);
      const uint32_t u = !(m > 0.0f) ? 0u : (m >= 255.0f ? 255u : uint32_t(m));
      out = (out << 8) | u;
    }
  }
  colourOut[idx] = out;
}

// ---- classification -----------------------------------------------------------

struct CubeGrid {
  uint32_t S[3];
  uint32_t M[3];
  uint32_t cubes;  // < 2^31 (checked on the host)
};

// Cube m = (x*My + y)*Mz + z (to3d, curves.h:22-27). The kernels walk cubes
// m0, m0 + kThreads, ... : the coordinates are divided out once, then stepped.
struct CubeCursor {
  uint32_t x, y, z;
  __device__ CubeCursor(const CubeGrid& g, uint32_t m) {
    z = m % g.M[2];
    const uint32_t xy = m / g.M[2];
    y = xy % g.M[1];
    x = xy / g.M[1];
  }
  __device__ void advance(const CubeGrid& g, uint32_t by) {
    z += by;
    while (z >= g.M[2]) {
      z -= g.M[2];
      if (++y == g.M[1]) {
        y = 0;
        ++x;
      }
    }
  }
};

__device__ inline uint64_t sampleIndex(const CubeGrid& g, uint32_t x, uint32_t y, uint32_t z) {
  return (uint64_t(x) * g.S[1] + y) * g.S[2] + z;
}

__device__ inline uint32_t cubeCase(const CubeGrid& g, const float* __restrict__ field,
                                    const CubeCursor& q, float iso) {
  const uint64_t sz = g.S[2], syz = uint64_t(g.S[1]) * g.S[2];
  const uint64_t b = sampleIndex(g, q.x, q.y, q.z);
  uint32_t ci = 0;
  const float v0 = field[b], v1 = field[b + syz], v2 = field[b + syz + sz], v3 = field[b + sz];
  const float v4 = field[b + 1], v5 = field[b + syz + 1], v6 = field[b + syz + sz + 1],
              v7 = field[b + sz + 1];
  ci |= (v0 < iso) ? 1u : 0u;
  ci |= (v1 < iso) ? 2u : 0u;
  ci |= (v2 < iso) ? 4u : 0u;
  ci |= (v3 < iso) ? 8u : 0u;
  ci |= (v4 < iso) ? 16u : 0u;
  ci |= (v5 < iso) ? 32u : 0u;
  ci |= (v6 < iso) ? 64u : 0u;
  ci |= (v7 < iso) ? 128u : 0u;
  return ci;
}

// mc_size (oclsph_kernel.h:545-596): triangles per block of kCubesPerBlock
// cubes. Cube m = block*1024 + j*256 + thread (j = 0..3). The case indices
// are kept (one byte per cube) when cases is not null.
// With tile flags of the density field that wrote the field: a block whose
// samples all lie in tiles written as zero has a uniform field, so every cube
// has 0 triangles (case 0 or 255) and the block writes a zero count without
// reading the field (its case bytes are not written: the emission never
// reads them for a block without triangles). (Testing the flags per cube
// instead of reading its 8 samples was slower: 88 -> 108 us at 256^3.)
struct TileFlags {
  const uint8_t* empty;  // null: no flags
  uint32_t size[3];      // samples per tile
  uint32_t tiles[3];     // tiles per axis
};
__global__ __launch_bounds__(kThreads) void k_mc_count(CubeGrid g, const float* __restrict__ field,
                                                       float iso, uint32_t* __restrict__ blockCount,
                                                       uint8_t* __restrict__ cases, TileFlags tf) {
  __shared__ uint32_t sm[kThreads / 64];
  if (tf.empty) {
    // the block's cubes [c0, c1]: samples x0..x1+1 and, within one x, y0..y1+1,
    // within one (x, y) z0..z1+1; else the whole y or z range
    const uint32_t c0 = blockIdx.x * kCubesPerBlock,
                   c1 = min(c0 + kCubesPerBlock, g.cubes) - 1u;
    const uint32_t x0 = c0 / (g.M[1] * g.M[2]), x1 = c1 / (g.M[1] * g.M[2]);
    const uint32_t y0 = (c0 / g.M[2]) % g.M[1], y1 = (c1 / g.M[2]) % g.M[1];
    const uint32_t sy0 = x0 == x1 ? y0 : 0u, sy1 = x0 == x1 ? y1 + 1u : g.S[1] - 1u;
    const bool oneColumn = x0 == x1 && y0 == y1;
    const uint32_t sz0 = oneColumn ? c0 % g.M[2] : 0u,
                   sz1 = oneColumn ? c1 % g.M[2] + 1u : g.S[2] - 1u;
    const uint32_t t0[3] = {x0 / tf.size[0], sy0 / tf.size[1], sz0 / tf.size[2]};
    const uint32_t n[3] = {(x1 + 1u) / tf.size[0] - t0[0] + 1u, sy1 / tf.size[1] - t0[1] + 1u,
                           sz1 / tf.size[2] - t0[2] + 1u};
    const uint32_t nt = n[0] * n[1] * n[2];
    bool empty = nt <= uint32_t(kThreads);  // more tiles than threads: no test
    if (empty && threadIdx.x < nt) {
      const uint32_t i = threadIdx.x;
      const uint32_t tx = t0[0] + i / (n[1] * n[2]), ty = t0[1] + (i / n[2]) % n[1],
                     tz = t0[2] + i % n[2];
      empty = tf.empty[(uint64_t(tx) * tf.tiles[1] + ty) * tf.tiles[2] + tz] != 0;
    }
    if (__syncthreads_and(empty)) {
      if (threadIdx.x == 0) blockCount[blockIdx.x] = 0u;
      return;
    }
  }
  uint32_t sum = 0;
  const uint32_t m0 = blockIdx.x * kCubesPerBlock + threadIdx.x;
  CubeCursor q(g, min(m0, g.cubes - 1u));
#pragma unroll
  for (int j = 0; j < kCubesPerThread; ++j) {
    const uint32_t m = m0 + uint32_t(j) * kThreads;
    if (j > 0) q.advance(g, kThreads);
    if (m < g.cubes) {
      const uint32_t ci = cubeCase(g, field, q, iso);
      if (cases) cases[m] = uint8_t(ci);
      sum += cTab.ntri[ci];
    }
  }
  // block sum
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
  if (__lane_id() == 0) sm[threadIdx.x >> 6] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < kThreads / 64; ++i) t += sm[i];
    blockCount[blockIdx.x] = t;
  }
}

// Exclusive scan of the per-block counts (one workgroup of 1024), total into
// total[0]: the host sum of oclsph.hpp:461-465, kept on the device.
__global__ __launch_bounds__(1024) void k_mc_scan(uint32_t* __restrict__ counts, uint32_t nb,
                                                  uint32_t* __restrict__ total) {
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
#pragma unroll 8
  for (uint32_t i = lo; i < hi; ++i) {
    const uint32_t t = counts[i];
    counts[i] = run;
    run += t;
  }
  if (threadIdx.x == 0) total[0] = tot;
}

// ---- emission -----------------------------------------------------------------

struct EmitArgs {
  CubeGrid g;
  float iso;
  float origin[3], step[3], scale;
  uint64_t cap;
  int attributes;
};

// Block-exclusive prefix of small counts (< 8) in thread order.
__device__ inline uint32_t blockPrefixSmall(uint32_t v, uint32_t* sm, uint32_t* total) {
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  const uint64_t lt = (uint64_t(1) << lane) - 1u;
  uint32_t pre = 0, wsum = 0;
#pragma unroll
  for (int bit = 0; bit < 3; ++bit) {
    const uint64_t bb = __ballot((v >> bit) & 1u);
    pre += uint32_t(__popcll(bb & lt)) << bit;
    wsum += uint32_t(__popcll(bb)) << bit;
  }
  if (lane == 0) sm[w] = wsum;
  __syncthreads();
  uint32_t wb = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kThreads / 64; ++i) {
    if (i < w) wb += sm[i];
    tot += sm[i];
  }
  __syncthreads();
  *total = tot;
  return wb + pre;
}

// One vertex of edge e of the cube at q: t = (iso - v_a)/(v_b - v_a),
// position mix(p_a, p_b, t) with p = (origin + o*step)*scale, and with
// attributes the normal and colour mixed the same way.
struct EdgeVertex {
  float p[3];
  float t;
  uint64_t ia, ib;  // sample indices of the edge's corners
};
__device__ inline EdgeVertex edgeVertex(const EmitArgs& A, const float* __restrict__ field,
                                        const CubeCursor& q, int e) {
  EdgeVertex r;
  const int a = cEdge[e][0], b = cEdge[e][1];
  const uint32_t oa[3] = {q.x + uint32_t(cCorner[a][0]), q.y + uint32_t(cCorner[a][1]),
                          q.z + uint32_t(cCorner[a][2])};
  const uint32_t ob[3] = {q.x + uint32_t(cCorner[b][0]), q.y + uint32_t(cCorner[b][1]),
                          q.z + uint32_t(cCorner[b][2])};
  r.ia = sampleIndex(A.g, oa[0], oa[1], oa[2]);
  r.ib = sampleIndex(A.g, ob[0], ob[1], ob[2]);
  const float va = field[r.ia], vb = field[r.ib];
  r.t = (A.iso - va) / (vb - va);
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float pa = __builtin_fmaf(float(oa[k]), A.step[k], A.origin[k]) * A.scale;
    const float pb = __builtin_fmaf(float(ob[k]), A.step[k], A.origin[k]) * A.scale;
    r.p[k] = __builtin_fmaf(pb - pa, r.t, pa);
  }
  return r;
}

// mc_eval (oclsph_kernel.h:599-735): edge intersections by linear
// interpolation t = (iso - v0)/(v1 - v0), mix(p0, p1, t); triangles of
// TriTable[ci] written at the scanned offset, in cube order. Blocks without
// triangles exit at once. The case indices come from the classify pass (one
// byte per cube), or, with kCases false, are classified again from the field
// (no case array). Each cube with triangles stamps its block-local index over
// its slots in LDS, then thread k emits the block's triangle k: the threads
// share the work evenly and consecutive threads write consecutive records.
constexpr uint32_t kMaxBlockTriangles = kCubesPerBlock * 5;  // at most 5 per cube

template <bool kCases>
__global__ __launch_bounds__(kThreads) void k_mc_emit(EmitArgs A, const float* __restrict__ field,
                                                      const uint8_t* __restrict__ cases,
                                                      const float4* __restrict__ normals,
                                                      const uint32_t* __restrict__ colours,
                                                      const uint32_t* __restrict__ blockOffset,
                                                      const uint32_t* __restrict__ total,
                                                      uint32_t nBlocks,
                                                      float* __restrict__ triPos,
                                                      float* __restrict__ triNrm,
                                                      uint32_t* __restrict__ triCol) {
  __shared__ uint32_t sm[kThreads / 64];
  __shared__ uint16_t owner[kMaxBlockTriangles];  // block-local cube of each triangle
  __shared__ uint8_t triInCube[kMaxBlockTriangles];  // its index within the cube
  __shared__ uint8_t caseOf[kCubesPerBlock];
  const uint64_t first = blockOffset[blockIdx.x];
  const uint64_t next = (blockIdx.x + 1 < nBlocks) ? blockOffset[blockIdx.x + 1] : total[0];
  if (next == first) return;  // uniform: no triangle in this block
  const uint32_t base = blockIdx.x * kCubesPerBlock;
  uint32_t run = 0;  // triangles of the block's earlier rounds
  for (uint32_t j = 0; j < uint32_t(kCubesPerThread); ++j) {
    const uint32_t ib = j * kThreads + threadIdx.x;  // block-local cube, cube order
    const uint32_t m = base + ib;
    uint32_t ci = 0;
    if (m < A.g.cubes) {
      if constexpr (kCases) {
        ci = cases[m];
      } else {
        ci = cubeCase(A.g, field, CubeCursor(A.g, m), A.iso);
      }
    }
    caseOf[ib] = uint8_t(ci);
    const uint32_t nt = cTab.ntri[ci];
    uint32_t tot;
    const uint32_t pre = run + blockPrefixSmall(nt, sm, &tot);
    for (uint32_t i = 0; i < nt; ++i) {
      owner[pre + i] = uint16_t(ib);
      triInCube[pre + i] = uint8_t(i);
    }
    run += tot;
  }
  __syncthreads();
  const uint32_t T = run;  // == next - first
  for (uint32_t k = threadIdx.x; k < T; k += kThreads) {
    const uint64_t slot = first + k;
    if (slot >= A.cap) break;
    const uint32_t ib = owner[k], i = triInCube[k];
    const uint32_t ci = caseOf[ib];
    const CubeCursor q(A.g, base + ib);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const int e = cTab.tri[ci][3 * i + v];
      const EdgeVertex r = edgeVertex(A, field, q, e);
      float* dst = triPos + slot * 9 + 3 * v;
      dst[0] = r.p[0];
      dst[1] = r.p[1];
      dst[2] = r.p[2];
      if (A.attributes) {
        float na[3] = {0.0f, 0.0f, 0.0f}, nb[3] = {0.0f, 0.0f, 0.0f};
        float ca[4], cb[4];
        if (normals) {
          const float4 n0 = normals[r.ia], n1 = normals[r.ib];
          na[0] = n0.x; na[1] = n0.y; na[2] = n0.z;
          nb[0] = n1.x; nb[1] = n1.y; nb[2] = n1.z;
        }
        if (colours) {
          const uint32_t c0 = colours[r.ia], c1 = colours[r.ib];
          for (int k2 = 0; k2 < 4; ++k2) {
            ca[k2] = float((c0 >> (24 - 8 * k2)) & 0xFFu);
            cb[k2] = float((c1 >> (24 - 8 * k2)) & 0xFFu);
          }
        } else {
          for (int k2 = 0; k2 < 4; ++k2) ca[k2] = cb[k2] = 255.0f;
        }
        float* nd = triNrm + slot * 9 + 3 * v;
        for (int k2 = 0; k2 < 3; ++k2) nd[k2] = __builtin_fmaf(nb[k2] - na[k2], r.t, na[k2]);
        uint32_t packed = 0;
        for (int k2 = 0; k2 < 4; ++k2) {
          const float mv = __builtin_fmaf(cb[k2] - ca[k2], r.t, ca[k2]);
          const uint32_t u = !(mv > 0.0f) ? 0u : (mv >= 255.0f ? 255u : uint32_t(mv));
          packed = (packed << 8) | u;
        }
        triCol[slot * 3 + v] = packed;
      }
    }
  }
}

}  // namespace

// The marching-cubes stream (default priority, like the main stream) and its
// events, created on first use: a context that never meshes keeps one stream.
static bool mcStreamReady(Ctx* c) {
  if (c->mcStream) return true;
  hipStream_t s = nullptr;
  hipEvent_t e[5] = {};
  bool ok = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess;
  for (int i = 0; ok && i < 5; ++i)
    ok = hipEventCreateWithFlags(&e[i], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    for (hipEvent_t x : e)
      if (x) (void)hipEventDestroy(x);
    if (s) (void)hipStreamDestroy(s);
    return false;
  }
  c->mcStream = s;
  c->evMcIn = e[0];
  c->evMcDone = e[1];
  c->pairGuard.ev = e[2];
  c->pairGuardAlt.ev = e[3];
  c->evPosFinal = e[4];
  c->posEventValid = false;  // nothing recorded yet
  return true;
}

// Scope in which launches go to the context's marching-cubes stream; outside
// a graph capture only (else, or if the stream cannot be made, they stay on
// the main stream). It first waits on the main stream for what the launches
// read there: the last write of P and the table (sampling), or, when the
// field was last written on the main stream, everything issued so far
// (classification, emission); else only for its own earlier work.
enum class McInput { Positions, Field };
struct OnMcStream {
  Ctx* c;
  bool on;
  OnMcStream(Ctx* ctx, McInput in) : c(ctx), on(!ctx->capturing && mcStreamReady(ctx)) {
    if (!on) return;
    hipError_t e = hipSuccess;
    if (in == McInput::Positions && c->posEventValid) {
      e = hipStreamWaitEvent(c->mcStream, c->evPosFinal, 0);
    } else if (in == McInput::Positions || c->fieldOnMain) {
      e = hipEventRecord(c->evMcIn, c->stream);
      if (e == hipSuccess) e = hipStreamWaitEvent(c->mcStream, c->evMcIn, 0);
    }
    if (e != hipSuccess) {
      // stay on the main stream, after the earlier work of the marching-cubes
      // stream (it may still read the field these launches rewrite)
      on = false;
      (void)hipStreamSynchronize(c->mcStream);
      c->mcBusy = false;
      return;
    }
    c->fieldOnMain = false;
    c->launchStream = c->mcStream;
  }
  ~OnMcStream() {
    if (!on) return;
    c->launchStream = nullptr;
    if (hipEventRecord(c->evMcDone, c->mcStream) == hipSuccess)
      c->mcBusy = true;
    else
      (void)hipStreamSynchronize(c->mcStream);
  }
};

int mcSample(Ctx* c, const lf_mc_params& m) {
  if (!c->gridValid) return fail(c, LF_ERR_STATE, "lf_mc_sample: no grid (call lf_grid_build or lf_pbf_step)");
  // the selected scene of a batch: its constants and its slice of the table
  const PbfConsts k = (c->batch && size_t(c->mcScene) < c->sceneConsts.size())
                          ? c->sceneConsts[c->mcScene]
                          : c->consts;
  const uint32_t* table = c->table.p + (c->batch ? k.keyBase : 0u);
  Lattice L{};
  if (m.field == LF_MC_FIELD_REFERENCE) {
    if (!(m.resolution > 0.0f)) return fail(c, LF_ERR_ARG, "lf_mc_sample: resolution must be > 0");
    for (int i = 0; i < 3; ++i) {  // oclsph.hpp:817-819
      L.S[i] = uint32_t(floorf(float(k.extent[i]) * m.resolution)) + 1u;
      L.origin[i] = k.minExt[i];
      L.step[i] = k.h / m.resolution;
    }
  } else if (m.field == LF_MC_FIELD_DENSITY) {
    for (int i = 0; i < 3; ++i) {
      if (m.samples[i] < 2) return fail(c, LF_ERR_ARG, "lf_mc_sample: samples must be >= 2 per axis");
      L.S[i] = m.samples[i];
      L.origin[i] = k.minB[i] / k.scale;
      const float span = (k.maxB[i] - k.minB[i]) / k.scale;
      L.step[i] = span / float(L.S[i] - 1u);
    }
  } else {
    return fail(c, LF_ERR_ARG, "lf_mc_sample: unknown field");
  }
  const uint64_t total = uint64_t(L.S[0]) * L.S[1] * L.S[2];
  if (total >= (uint64_t(1) << 40)) return fail(c, LF_ERR_ARG, "lf_mc_sample: lattice too large");
  int rc = ensure(c, c->field, size_t(total), "mc field");
  if (rc) return rc;
  for (int i = 0; i < 3; ++i) {
    c->S[i] = L.S[i];
    c->mcOrigin[i] = L.origin[i];
    c->mcStep[i] = L.step[i];
  }
  const unsigned nb = unsigned((total + kThreads - 1) / kThreads);
  if (m.field == LF_MC_FIELD_REFERENCE) {
    // on the main stream (it reads the state, which the next step rewrites),
    // after the last classification and emission have read the field
    if (mcJoin(c)) return fail(c, LF_ERR_HIP, "lf_mc_sample: stream wait");
    if ((rc = ensure(c, c->mcNormal, size_t(total), "mc normals"))) return rc;
    if ((rc = ensure(c, c->mcColour, size_t(total), "mc colours"))) return rc;
    rc = launch(c, "mc_reference_field", k_mc_reference, dim3(nb), dim3(kThreads), 0, total, L, k,
                m.resolution, m.particle_size, m.particle_influence, table,
                (const uint8_t*)c->T[c->cur].p, (const float4*)c->X[c->cur].p,
                (const float4*)c->V[c->cur].p, c->field.p, c->mcNormal.p, c->mcColour.p,
                (const uint32_t*)c->health.p);
    c->fieldOnMain = true;
    c->mcTileFlags = false;
    c->mcAttr = 1;
  } else {
    // tile shapes (lf_exec_params.mc_tile): columns kWX x kWY per wave,
    // kNZ samples per column, waves stacked in z
    struct Shape {
      uint32_t tx, ty, tz, threads;
    };
    static const Shape kShapes[kMcTiles] = {{8, 8, 16, 256}, {4, 16, 16, 256}, {8, 8, 8, 256},
                                            {8, 8, 4, 256}};
    const Shape sh = kShapes[c->mcTile < kMcTiles ? c->mcTile : 0];
    const uint32_t tX = (L.S[0] + sh.tx - 1) / sh.tx, tY = (L.S[1] + sh.ty - 1) / sh.ty,
                   tZ = (L.S[2] + sh.tz - 1) / sh.tz;
    const uint64_t tiles = uint64_t(tX) * tY * tZ;
    if (tiles >= (uint64_t(1) << 31)) return fail(c, LF_ERR_ARG, "lf_mc_sample: lattice too large");
    // on the marching-cubes stream, after the step: it reads only P and the
    // table, which the next step leaves alone (it writes the other pair)
    OnMcStream onMc(c, McInput::Positions);
    // 2,048 staged particles (32 KB): the halos of the 1M dam break's 256^3
    // tiles hold at most 1,932 (8 x 8 x 16; median 1,663); a tile above it
    // reads global memory
    auto kern = c->mcTile == 1   ? k_mc_density<4, 16, 4, 4, 2048>
                : c->mcTile == 2 ? k_mc_density<8, 8, 2, 4, 2048>
                : c->mcTile == 3 ? k_mc_density<8, 8, 1, 4, 2048>
                                 : k_mc_density<8, 8, 4, 4, 2048>;
    if ((rc = ensure(c, c->mcTileEmpty, size_t(tiles), "mc tile flags"))) return rc;
    rc = launch(c, "mc_density_field", kern, dim3(unsigned(tiles)), dim3(sh.threads), 0, L, k, table,
                (const float4*)c->P.p, c->field.p, tY, tZ, (const uint32_t*)c->health.p,
                c->mcTileEmpty.p);
    // a diverged state returns before writing field or flags: no flags then
    c->mcTileFlags = rc == LF_OK;
    c->mcTileSize[0] = sh.tx;
    c->mcTileSize[1] = sh.ty;
    c->mcTileSize[2] = sh.tz;
    c->mcTiles[0] = tX;
    c->mcTiles[1] = tY;
    c->mcTiles[2] = tZ;
    c->lastDensity = m;
    c->haveDensity = true;
    if (rc == LF_OK && onMc.on) {
      if (hipEventRecord(c->pairGuard.ev, c->mcStream) == hipSuccess)
        c->pairGuard.pending = true;
      else
        (void)hipStreamSynchronize(c->mcStream);
    }
    c->mcAttr = 0;
  }
  if (rc) return rc;
  c->mcScale = k.scale;
  c->mcSampled = true;
  c->mcCounted = c->mcEmitted = false;
  return LF_OK;
}

// Cube indices are 32-bit in the kernels (with a block of headroom).
static bool cubesFit(const Ctx* c) {
  return uint64_t(c->S[0] - 1) * (c->S[1] - 1) * (c->S[2] - 1) < (uint64_t(1) << 31);
}

static CubeGrid cubeGrid(const Ctx* c) {
  CubeGrid g{};
  for (int i = 0; i < 3; ++i) {
    g.S[i] = c->S[i];
    g.M[i] = c->S[i] - 1u;
  }
  g.cubes = uint32_t(uint64_t(g.M[0]) * g.M[1] * g.M[2]);
  return g;
}

int mcClassifyCount(Ctx* c, const lf_mc_params& m) {
  if (!c->mcSampled) return fail(c, LF_ERR_STATE, "lf_mc_classify_count: no field (call lf_mc_sample)");
  if (!cubesFit(c)) return fail(c, LF_ERR_ARG, "marching cubes: more than 2^31 cubes");
  OnMcStream onMc(c, McInput::Field);
  const CubeGrid g = cubeGrid(c);
  const uint64_t nbl = (uint64_t(g.cubes) + kCubesPerBlock - 1) / kCubesPerBlock;
  int rc = ensure(c, c->mcBlock, size_t(nbl) + 1, "mc block counts");
  if (rc) return rc;
  if ((rc = ensure(c, c->mcTotal, 4, "mc total"))) return rc;
  if ((rc = ensure(c, c->mcCases, std::max<size_t>(g.cubes, 1), "mc cases"))) return rc;
  if (nbl > 0) {
    TileFlags tf{};
    if (c->mcTileFlags && !c->keepCases) {
      tf.empty = c->mcTileEmpty.p;
      for (int i = 0; i < 3; ++i) {
        tf.size[i] = c->mcTileSize[i];
        tf.tiles[i] = c->mcTiles[i];
      }
    }
    rc = launch(c, "mc_classify_count", k_mc_count, dim3(unsigned(nbl)), dim3(kThreads), 0, g,
                (const float*)c->field.p, m.isolevel, c->mcBlock.p, c->mcCases.p, tf);
    if (rc) return rc;
  }
  c->mcIso = m.isolevel;
  rc = launch(c, "mc_scan", k_mc_scan, dim3(1), dim3(1024), 0, c->mcBlock.p, uint32_t(nbl),
              c->mcTotal.p);
  if (rc) return rc;
  c->mcCounted = true;
  return LF_OK;
}

int mcEmit(Ctx* c, const lf_mc_params& m) {
  if (!c->mcCounted) return fail(c, LF_ERR_STATE, "lf_mc_emit: no counts (call lf_mc_classify_count)");
  if (!(m.isolevel == c->mcIso))
    return fail(c, LF_ERR_ARG, "lf_mc_emit: isolevel differs from the one lf_mc_classify_count used");
  OnMcStream onMc(c, McInput::Field);
  const CubeGrid g = cubeGrid(c);
  const uint64_t nbl = (uint64_t(g.cubes) + kCubesPerBlock - 1) / kCubesPerBlock;
  int rc = ensure(c, c->triPos, c->triCap * 9, "triangle positions");
  if (rc) return rc;
  const int attr = (m.attributes != 0) ? 1 : 0;
  if (attr) {
    if ((rc = ensure(c, c->triNrm, c->triCap * 9, "triangle normals"))) return rc;
    if ((rc = ensure(c, c->triCol, c->triCap * 3, "triangle colours"))) return rc;
  }
  EmitArgs A{};
  A.g = g;
  A.iso = m.isolevel;
  for (int i = 0; i < 3; ++i) {
    A.origin[i] = c->mcOrigin[i];
    A.step[i] = c->mcStep[i];
  }
  A.scale = c->mcScale;
  A.cap = c->triCap;
  A.attributes = attr;
  if (nbl > 0) {
    rc = launch(c, "mc_emit", k_mc_emit<true>, dim3(unsigned(nbl)), dim3(kThreads), 0, A,
                (const float*)c->field.p, (const uint8_t*)c->mcCases.p,
                (const float4*)((attr && c->mcAttr) ? c->mcNormal.p : nullptr),
                (const uint32_t*)((attr && c->mcAttr) ? c->mcColour.p : nullptr),
                (const uint32_t*)c->mcBlock.p, (const uint32_t*)c->mcTotal.p, uint32_t(nbl),
                c->triPos.p, attr ? c->triNrm.p : (float*)nullptr,
                attr ? c->triCol.p : (uint32_t*)nullptr);
    if (rc) return rc;
  }
  c->mcEmitted = true;
  return LF_OK;
}

// ---- meshing without allocation (lf_mc_mesh_static) ---------------------------

// The per-block triangle counts (then offsets) of lf_mc_mesh_static: a
// file-scope device array, so the call allocates nothing. Each GPU holds its
// own copy until the process ends; every call on that GPU shares it, and
// kernel arguments take the address hipGetSymbolAddress returned on that
// device (the array's name is a host-side stand-in). The count pass writes
// every entry it uses, so nothing has to start at zero.
constexpr uint32_t kStaticMaxBlocks = 1u << 18;  // 2^28 cubes, 1 MiB
static __device__ __align__(256) uint32_t scratch[kStaticMaxBlocks];

constexpr int kMaxDevices = 64;
static std::mutex gScratchMutex;
static uint32_t* gScratch[kMaxDevices];  // the array's address on each device

int mcStaticInit(int device, uint32_t** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return LF_ERR_ARG;
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (!gScratch[device]) {
    void* p = nullptr;  // device is current
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LF_ERR_HIP;
    gScratch[device] = static_cast<uint32_t*>(p);
  }
  *out = gScratch[device];
  return LF_OK;
}

int mcMeshStatic(Ctx* c, const uint32_t S[3], const float origin[3], const float step[3],
                 float scale, float iso, const float* field, float* triPos, size_t triCap,
                 uint32_t* ntri) {
  if (!c->mcScratch) return fail(c, LF_ERR_STATE, "lf_mc_mesh_static: no block-count array on this device");
  uint64_t cubes = 1;
  CubeGrid g{};
  for (int i = 0; i < 3; ++i) {
    if (S[i] < 2) return fail(c, LF_ERR_ARG, "lf_mc_mesh_static: samples must be >= 2 per axis");
    g.S[i] = S[i];
    g.M[i] = S[i] - 1u;
    cubes *= g.M[i];
  }
  const uint64_t nbl = (cubes + kCubesPerBlock - 1) / kCubesPerBlock;
  if (nbl > kStaticMaxBlocks)
    return fail(c, LF_ERR_ARG, "lf_mc_mesh_static: " + std::to_string(cubes) +
                                   " cubes; at most " +
                                   std::to_string(uint64_t(kStaticMaxBlocks) * kCubesPerBlock));
  g.cubes = uint32_t(cubes);
  EmitArgs A{};
  A.g = g;
  A.iso = iso;
  for (int i = 0; i < 3; ++i) {
    A.origin[i] = origin[i];
    A.step[i] = step[i];
  }
  A.scale = scale;
  A.cap = triCap;
  A.attributes = 0;
  // plain launches on the context's stream: no timing events, no buffers
  hipLaunchKernelGGL(k_mc_count, dim3(unsigned(nbl)), dim3(kThreads), 0, c->stream, g, field, iso,
                     c->mcScratch, (uint8_t*)nullptr, TileFlags{});
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hipFail(c, e, "mc_classify_count");
  hipLaunchKernelGGL(k_mc_scan, dim3(1), dim3(1024), 0, c->stream, c->mcScratch, uint32_t(nbl), ntri);
  if ((e = hipGetLastError()) != hipSuccess) return hipFail(c, e, "mc_scan");
  if (triCap == 0) return LF_OK;
  hipLaunchKernelGGL(k_mc_emit<false>, dim3(unsigned(nbl)), dim3(kThreads), 0, c->stream, A, field,
                     (const uint8_t*)nullptr, (const float4*)nullptr, (const uint32_t*)nullptr,
                     (const uint32_t*)c->mcScratch, (const uint32_t*)ntri, uint32_t(nbl), triPos,
                     (float*)nullptr, (uint32_t*)nullptr);
  if ((e = hipGetLastError()) != hipSuccess) return hipFail(c, e, "mc_emit");
  return LF_OK;
}

}  // namespace lf
