// Mesh colliders (oclsph_collision.h:7-58, used by sph_delta at
// oclsph_kernel.h:364-368): the per-triangle constants of collideTriangle2 and
// a uniform grid over the scenes' bounds that lists, per cell, the triangles
// that can hit a point of the cell.
//
// A triangle hits p when its plane lies within 40 world units of p and the
// projection of p onto the plane falls inside it, i.e. p lies in the
// triangle's prism: the triangle swept +-40 along its normal. A cell lists the
// triangle unless a separating plane shows that the cell and the prism are
// disjoint: the two slab faces and the three side faces of the prism, each
// moved outward by a margin far above the rounding of the FP32 test (so a
// triangle left out could not pass that test for any point of the cell). Lists
// keep triangle order, so the first hit among a cell's triangles is the first
// hit of the whole list, as in the reference's loop. Triangles whose FP32
// test is badly conditioned (angle at the first corner under ~1.8 degrees)
// skip the side faces; degenerate ones (no normal) never pass the test and
// are listed nowhere.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "lf_internal.hpp"

namespace lf {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kGridCells = 64;           // cells along the longest side of the bounds
constexpr uint32_t kGridMinTriangles = 8;     // fewer triangles: scan them all
constexpr uint64_t kGridMaxEntries = 1u << 26; // above this, scan (memory bound on the lists)
constexpr int kBoundDoubles = 32;

// Per triangle, in doubles: nn[3], dn, then for each edge k an outward
// in-plane unit normal m_k[3] and its offset o_k = m_k . (edge start), the
// slab margin, the side margin, a flag (0: never hits, 1: full test, 2: slab
// test only) and the three corners.
enum : int {
  kNN = 0, kDN = 3, kM0 = 4, kSlabMargin = 16, kSideMargin = 17, kFlag = 18, kV = 19
};

struct GridDesc {
  double origin[3];
  double cell;
  double half;  // half cell size, widened for the rounding of the cell lookup
  uint32_t dims[3];
  uint32_t cells;
};

__device__ inline bool cellMayHit(const double* __restrict__ B, const double c[3], double h) {
  const double flag = B[kFlag];
  if (flag == 0.0) return false;
  // slab |nn . p - dn| <= 40
  const double s = B[kNN] * c[0] + B[kNN + 1] * c[1] + B[kNN + 2] * c[2] - B[kDN];
  const double r = h * (fabs(B[kNN]) + fabs(B[kNN + 1]) + fabs(B[kNN + 2]));
  if (fabs(s) > 40.0 + r + B[kSlabMargin]) return false;
  if (flag == 2.0) return true;
  // side faces
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double* m = B + kM0 + 4 * k;
    const double d = m[0] * c[0] + m[1] * c[1] + m[2] * c[2] - m[3];
    const double rk = h * (fabs(m[0]) + fabs(m[1]) + fabs(m[2]));
    if (d - rk > B[kSideMargin]) return false;
  }
  // nn x (unit axis): with the side normals these are all the separating
  // directions of the prism's cross-section and the cell's outline seen
  // along nn (2-D separating axes of a triangle and a hexagon)
  const double* n = B + kNN;
  const double* V = B + kV;
  const double axes[3][3] = {{0.0, n[2], -n[1]}, {-n[2], 0.0, n[0]}, {n[1], -n[0], 0.0}};
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double* a = axes[j];
    const double len = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    if (!(len > 1e-6)) continue;  // nn along this axis: the side faces cover it
    const double p0 = a[0] * V[0] + a[1] * V[1] + a[2] * V[2];
    const double p1 = a[0] * V[3] + a[1] * V[4] + a[2] * V[5];
    const double p2 = a[0] * V[6] + a[1] * V[7] + a[2] * V[8];
    const double lo = fmin(p0, fmin(p1, p2)), hi = fmax(p0, fmax(p1, p2));
    const double pc = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
    const double r = h * (fabs(a[0]) + fabs(a[1]) + fabs(a[2]));
    const double m = B[kSideMargin] * len;
    if (pc - r > hi + m || pc + r < lo - m) return false;
  }
  return true;
}

// One wave per cell, lanes over triangles in index order. FILL = false
// counts, FILL = true writes the ascending list at cellStart[cell].
template <bool FILL>
__global__ __launch_bounds__(kThreads) void k_collider_cells(GridDesc d,
                                                             const double* __restrict__ bounds,
                                                             uint32_t n,
                                                             uint32_t* __restrict__ counts,
                                                             const uint32_t* __restrict__ cellStart,
                                                             uint32_t* __restrict__ cellTris) {
  const uint32_t cell = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6);
  if (cell >= d.cells) return;  // whole wave
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t ix = cell % d.dims[0], iy = (cell / d.dims[0]) % d.dims[1],
                 iz = cell / (d.dims[0] * d.dims[1]);
  const double c[3] = {d.origin[0] + (double(ix) + 0.5) * d.cell,
                       d.origin[1] + (double(iy) + 0.5) * d.cell,
                       d.origin[2] + (double(iz) + 0.5) * d.cell};
  uint32_t run = FILL ? cellStart[cell] : 0u;
  for (uint32_t j0 = 0; j0 < n; j0 += 64) {
    const uint32_t j = j0 + lane;
    const bool hit = j < n && cellMayHit(bounds + size_t(j) * kBoundDoubles, c, d.half);
    const uint64_t bb = __ballot(hit);
    if (FILL && hit) cellTris[run + uint32_t(__popcll(bb & ((uint64_t(1) << lane) - 1u)))] = j;
    run += uint32_t(__popcll(bb));
  }
  if (!FILL && lane == 0) counts[cell] = run;
}

double dot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

}  // namespace

// The float constants collideTriangle2 derives from a triangle
// (oclsph_collision.h:7-22, 25-33), folded once per triangle.
int setColliders(Ctx* c, size_t n, const lf_triangle* tris) {
  // the same mesh again (a caller that passes its colliders every step):
  // keep the constants and the grid
  if (n == c->hostColliders.size() && n == c->nColliders &&
      (n == 0 || std::memcmp(tris, c->hostColliders.data(), n * sizeof(lf_triangle)) == 0))
    return LF_OK;
  std::vector<ColliderTri> host(n);
  for (size_t i = 0; i < n; ++i) {
    const lf_triangle& t = tris[i];
    ColliderTri& T = host[i];
    float e1[3], e2[3], nv[3];
    for (int k = 0; k < 3; ++k) {
      T.a[k] = t.v0[k];
      e1[k] = t.v1[k] - t.v0[k];  // b - a
      e2[k] = t.v2[k] - t.v0[k];  // c - a
    }
    nv[0] = e1[1] * e2[2] - e1[2] * e2[1];
    nv[1] = e1[2] * e2[0] - e1[0] * e2[2];
    nv[2] = e1[0] * e2[1] - e1[1] * e2[0];
    const float len = std::sqrt(dot3(nv, nv));
    for (int k = 0; k < 3; ++k) {
      T.nn[k] = nv[k] / len;
      T.v0[k] = e2[k];
      T.v1[k] = e1[k];
    }
    T.dn = dot3(T.nn, T.a);
    T.dot00 = dot3(T.v0, T.v0);
    T.dot01 = dot3(T.v0, T.v1);
    T.dot11 = dot3(T.v1, T.v1);
    T.invDenom = 1.0f / (T.dot00 * T.dot11 - T.dot01 * T.dot01);
    T.pad[0] = T.pad[1] = T.pad[2] = 0.0f;
  }
  if (n > 0) {
    int rc = ensure(c, c->colliders, n, "colliders");
    if (rc) return rc;
    LF_HIP(c, hipMemcpyAsync(c->colliders.p, host.data(), n * sizeof(ColliderTri),
                             hipMemcpyHostToDevice, c->stream));
    LF_HIP(c, hipStreamSynchronize(c->stream));
  }
  dropGraph(c);  // captured delta launches hold the old colliders
  c->hostColliders.assign(tris, tris + n);
  c->nColliders = uint32_t(n);
  c->colliderGridValid = false;
  return LF_OK;
}

// Builds (or keeps) the grid for the current scenes' bounds; called before
// each delta pass. Fills c->colliderGrid.
int prepareColliderGrid(Ctx* c) {
  const uint32_t n = c->nColliders;
  ColliderGrid& g = c->colliderGrid;
  if (n == 0) {
    g = ColliderGrid{};
    return LF_OK;
  }
  float box[6] = {1e30f, 1e30f, 1e30f, -1e30f, -1e30f, -1e30f};
  for (const PbfConsts& k : c->sceneConsts)
    for (int a = 0; a < 3; ++a) {
      box[a] = std::min(box[a], k.minB[a]);
      box[3 + a] = std::max(box[3 + a], k.maxB[a]);
    }
  if (c->colliderGridValid && std::memcmp(box, c->colliderGridBox, sizeof(box)) == 0 &&
      g.tris == c->colliders.p)
    return LF_OK;
  if (c->capturing)
    return fail(c, LF_ERR_STATE, "colliders or bounds changed during graph capture");
  dropGraph(c);
  g = ColliderGrid{};
  g.tris = c->colliders.p;
  g.n = n;
  g.useGrid = 0;
  std::memcpy(c->colliderGridBox, box, sizeof(box));
  c->colliderGridValid = true;
  double E[3], maxE = 0.0;
  bool finite = true;
  for (int a = 0; a < 3; ++a) {
    E[a] = double(box[3 + a]) - double(box[a]);
    finite = finite && std::isfinite(E[a]) && E[a] >= 0.0;
    maxE = std::max(maxE, E[a]);
  }
  if (n < kGridMinTriangles || !finite || !(maxE > 0.0)) return LF_OK;  // scan them all
  GridDesc d{};
  d.cell = maxE / double(kGridCells);
  double M = 40.0;  // largest coordinate magnitude the tests meet
  for (int a = 0; a < 3; ++a) {
    d.dims[a] = uint32_t(std::ceil(E[a] / d.cell)) + 2u;  // one cell of padding per side
    d.origin[a] = double(box[a]) - d.cell;
    M = std::max(M, std::fabs(d.origin[a]) + (d.dims[a] + 1) * d.cell);
  }
  for (const lf_triangle& t : c->hostColliders)
    for (int k = 0; k < 3; ++k)
      M = std::max({M, double(std::fabs(t.v0[k])), double(std::fabs(t.v1[k])),
                    double(std::fabs(t.v2[k]))});
  d.cells = d.dims[0] * d.dims[1] * d.dims[2];
  // the float cell lookup of a point may land one rounding step off a face
  d.half = 0.5 * d.cell + 1e-5 * (M + d.cell);
  // inclusion records (double), from the same float constants the test uses
  std::vector<ColliderTri> tris(n);
  LF_HIP(c, hipMemcpyAsync(tris.data(), c->colliders.p, n * sizeof(ColliderTri),
                           hipMemcpyDeviceToHost, c->stream));
  LF_HIP(c, hipStreamSynchronize(c->stream));
  std::vector<double> B(size_t(n) * kBoundDoubles, 0.0);
  for (uint32_t i = 0; i < n; ++i) {
    const ColliderTri& T = tris[i];
    const lf_triangle& t = c->hostColliders[i];
    double* b = &B[size_t(i) * kBoundDoubles];
    bool ok = std::isfinite(T.dn) && std::isfinite(T.invDenom);
    for (int k = 0; k < 3; ++k) ok = ok && std::isfinite(T.nn[k]);
    if (!ok) continue;  // flag 0: the FP32 test never passes
    for (int k = 0; k < 3; ++k) b[kNN + k] = T.nn[k];
    b[kDN] = T.dn;
    const double V[3][3] = {{t.v0[0], t.v0[1], t.v0[2]},
                            {t.v1[0], t.v1[1], t.v1[2]},
                            {t.v2[0], t.v2[1], t.v2[2]}};
    const double nn[3] = {T.nn[0], T.nn[1], T.nn[2]};
    double maxEdge = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double* s = V[k];
      const double* e = V[(k + 1) % 3];
      const double* o = V[(k + 2) % 3];
      const double ed[3] = {e[0] - s[0], e[1] - s[1], e[2] - s[2]};
      maxEdge = std::max(maxEdge, std::sqrt(dot(ed, ed)));
      double m[3] = {ed[1] * nn[2] - ed[2] * nn[1], ed[2] * nn[0] - ed[0] * nn[2],
                     ed[0] * nn[1] - ed[1] * nn[0]};
      const double ml = std::sqrt(dot(m, m));
      if (!(ml > 0.0)) {
        ok = false;
        break;
      }
      for (double& x : m) x /= ml;
      const double toOther[3] = {o[0] - s[0], o[1] - s[1], o[2] - s[2]};
      if (dot(m, toOther) > 0.0)
        for (double& x : m) x = -x;
      double* mk = b + kM0 + 4 * k;
      mk[0] = m[0];
      mk[1] = m[1];
      mk[2] = m[2];
      mk[3] = dot(m, s);
    }
    for (int k = 0; k < 3; ++k)
      for (int a = 0; a < 3; ++a) b[kV + 3 * k + a] = V[k][a];
    // conditioning of the barycentric solve: 1 / sin^2 of the corner angle
    const double d00 = T.dot00, d01 = T.dot01, d11 = T.dot11;
    const double det = d00 * d11 - d01 * d01;
    const double kappa = det > 0.0 ? (d00 * d11) / det : 1e300;
    // FP32 rounding of the test, bounded generously: t = dn - nn.p errs by a
    // few ulps of |dn| + |p|; the projected point by a few ulps of |p| + 40
    // and the barycentric solve by ~2 eps kappa times the edge length
    b[kSlabMargin] = 1e-5 * (1.0 + std::fabs(double(T.dn)) + 3.0 * M);
    b[kSideMargin] = 2e-4 * M + 3e-3 * maxEdge + 1e-6;
    b[kFlag] = (!ok || !(kappa <= 1e3)) ? 2.0 : 1.0;
  }
  int rc;
  if ((rc = ensure(c, c->colliderBounds, B.size(), "collider bounds"))) return rc;
  if ((rc = ensure(c, c->colliderCounts, d.cells, "collider grid counts"))) return rc;
  if ((rc = ensure(c, c->colliderCellStart, size_t(d.cells) + 1, "collider grid"))) return rc;
  LF_HIP(c, hipMemcpyAsync(c->colliderBounds.p, B.data(), B.size() * sizeof(double),
                           hipMemcpyHostToDevice, c->stream));
  const unsigned blocks = unsigned((d.cells + kThreads / 64 - 1) / (kThreads / 64));
  rc = launch(c, "collider_grid_count", k_collider_cells<false>, dim3(blocks), dim3(kThreads), 0, d,
              (const double*)c->colliderBounds.p, n, c->colliderCounts.p,
              (const uint32_t*)nullptr, (uint32_t*)nullptr);
  if (rc) return rc;
  std::vector<uint32_t> counts(d.cells), start(size_t(d.cells) + 1);
  LF_HIP(c, hipMemcpyAsync(counts.data(), c->colliderCounts.p, d.cells * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, c->stream));
  LF_HIP(c, hipStreamSynchronize(c->stream));
  uint64_t total = 0;
  for (uint32_t i = 0; i < d.cells; ++i) {
    start[i] = uint32_t(std::min<uint64_t>(total, 0xFFFFFFFFu));
    total += counts[i];
  }
  if (total > kGridMaxEntries) return LF_OK;  // lists too long: scan
  start[d.cells] = uint32_t(total);
  if ((rc = ensure(c, c->colliderCellTris, std::max<size_t>(total, 1), "collider grid lists")))
    return rc;
  LF_HIP(c, hipMemcpyAsync(c->colliderCellStart.p, start.data(), start.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, c->stream));
  rc = launch(c, "collider_grid_fill", k_collider_cells<true>, dim3(blocks), dim3(kThreads), 0, d,
              (const double*)c->colliderBounds.p, n, (uint32_t*)nullptr,
              (const uint32_t*)c->colliderCellStart.p, c->colliderCellTris.p);
  if (rc) return rc;
  LF_HIP(c, hipStreamSynchronize(c->stream));
  g.cellStart = c->colliderCellStart.p;
  g.cellTris = c->colliderCellTris.p;
  for (int a = 0; a < 3; ++a) {
    g.origin[a] = float(d.origin[a]);
    g.dims[a] = d.dims[a];
  }
  g.invCell = float(1.0 / d.cell);
  g.useGrid = 1;
  return LF_OK;
}

}  // namespace lf
