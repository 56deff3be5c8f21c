/*
 * ORACLE — test infrastructure, not product code.
 *
 * A CPU (C++/OpenMP) restatement of libfluid's Position Based Fluids step and
 * its marching-cubes surface extraction, templated on the arithmetic type:
 * F = float gives the FP32 "reference semantics" oracle (also the timed CPU
 * baseline), F = double the FP64 accuracy oracle. Only tests/, the smoke check
 * and the CPU-baseline leg of bench.py use it, always as the checker.
 *
 * Paths cited are relative to the reference root (tom91136/libfluid).
 *
 * Arithmetic contract (shared with the HIP kernels, which are written
 * separately against the same contract):
 *  - built with -ffp-contract=off; every fused multiply-add is an explicit
 *    std::fma, written where the reference writes mad() or where the
 *    contract below says so;
 *  - sqrt and division are IEEE correctly rounded (the reference's
 *    fast_distance / native_divide / unsafe-math builds are implementation
 *    defined, clutils.hpp:164-169, so they cannot be pinned);
 *  - squared distance d2 = fma(dz,dz, fma(dy,dy, dx*dx)), d = a - b;
 *  - W6(r)  = r <= h ? poly6 * t*t*t, t = fma(-r, r, h*h)  : 0
 *    gradW  = 1e-8 <= r <= h ? d * (spiky * ((h-r)*(h-r) / r)) : 0
 *    (oclsph_kernel.h:50-71);
 *  - divisions by constants are multiplications by the folded reciprocal:
 *    W6(r) / W6(dq) and the / rho0 of sph_delta (oclsph_kernel.h:354-358);
 *  - XSPH and the density field take W6 on d2: t = h*h - d2;
 *  - neighbour candidates: the 27 cells around the particle's own cell, x
 *    fastest then y then z, each cell's particles in sorted order, no radius
 *    cut (oclsph_kernel.h:150-192);
 *  - Jacobi delta pass (the reference updates p* in place while other
 *    work-items read it, oclsph_kernel.h:355,374: a data race);
 *  - stable sort by (min(key, G), caller index): the reference's std::sort
 *    is unstable (oclsph.hpp:735-738), and keys >= G belong to no cell
 *    (oclsph_kernel.h:186-187), so they share one last bucket in caller order.
 *
 * Parity pins (tests/test_oracle.py): Morton encode/decode against the
 * reference's own curves.h built in oracle/_ref and against a bit-loop
 * interleave; cell table against a brute-force lower bound; the 27-cell
 * candidate sets against an all-pairs |dcell|_inf <= 1 search; W6 against its
 * closed-form integral; lambda against the closed lattice sum (rho = 8078.2013,
 * lambda = -4.4429e-4 at s=0.05, h=0.1); delta = 0 for a symmetric particle;
 * marching-cubes tables against the reference's digests and oracle/_ref; a
 * sphere field's mesh area against 4 pi r^2.
 */
#ifndef LF_PBF_ORACLE_HPP
#define LF_PBF_ORACLE_HPP

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "fluid/mctable.hpp"
#include "libfluid_amd.h"

namespace oracle {

/* curves.h:56-72: spread 10 bits per axis, x at bit 3i, y at 3i+1, z at 3i+2. */
inline uint32_t spread10(uint32_t v) {
  v &= 1023u;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
inline uint32_t morton(uint32_t x, uint32_t y, uint32_t z) {
  return spread10(x) | (spread10(y) << 1) | (spread10(z) << 2);
}
/* curves.h:29-54 */
inline uint32_t compact10(uint32_t v) {
  v &= 0x09249249u;
  v = (v | (v >> 2)) & 0x030C30C3u;
  v = (v | (v >> 4)) & 0x0300F00Fu;
  v = (v | (v >> 8)) & 0x030000FFu;
  v = (v | (v >> 16)) & 0x000003FFu;
  return v;
}
inline void demorton(uint32_t k, uint32_t& x, uint32_t& y, uint32_t& z) {
  x = compact10(k);
  y = compact10(k >> 1);
  z = compact10(k >> 2);
}

/* FP32 pi as the OpenCL M_PI_F the kernels fold their constants with. */
template <typename F> inline F pi() { return F(3.14159265358979323846); }

/* Grid-cell coordinate of a solver-space offset q = (p - minExt) / h:
 * static_cast<size_t> of oclsph.hpp:327-332 then the 10-bit mask of
 * curves.h:56-72. Negative / NaN offsets (undefined in the reference) map to 0. */
template <typename F> inline uint32_t cellCoord(F q) {
  if (!(q >= F(0))) return 0u;
  if (q >= F(4294967296.0)) return 1023u;
  return static_cast<uint32_t>(q) & 1023u;
}

template <typename F> struct Consts {
  F h, hh, poly6, spiky, rhoRecip, wdq, wdqRecip, invDt, minExt[3], maxExt[3], minB[3],
      maxB[3], scale, dt, rho0, cfm, corrK, vd, xsph, xsphRho;
  uint32_t extent[3], G, iterations, flags;
  explicit Consts(const lf_pbf_params& p) {
    h = F(p.h);
    hh = h * h;
    const F hhh = hh * h;
    /* oclsph_kernel.h:50-51, folded in F: 315/(64 pi h^9), -45/(pi h^6). */
    poly6 = F(315) / ((F(64) * pi<F>()) * ((hhh * hhh) * hhh));
    spiky = -(F(45) / ((pi<F>() * hhh) * hhh));
    rho0 = F(p.rest_density);
    rhoRecip = F(1) / rho0;
    const F dq = F(p.corr_dq) * h;
    wdq = W6(dq);
    wdqRecip = F(1) / wdq;
    dt = F(p.dt);
    invDt = F(1) / dt;
    scale = F(p.scale);
    cfm = F(p.cfm_epsilon);
    corrK = F(p.corr_k);
    vd = F(p.vd);
    xsph = F(p.xsph_c);
    xsphRho = xsph / rho0;
    iterations = p.iterations;
    flags = p.flags;
    const F pad = h * F(2); /* oclsph.hpp:340 */
    for (int k = 0; k < 3; ++k) {
      minB[k] = F(p.min_bound[k]);
      maxB[k] = F(p.max_bound[k]);
      minExt[k] = (minB[k] / scale) - pad; /* oclsph.hpp:359-360 */
      maxExt[k] = (maxB[k] / scale) + pad;
      extent[k] = static_cast<uint32_t>((maxExt[k] - minExt[k]) / h); /* :386 */
    }
    G = morton(extent[0], extent[1], extent[2]); /* oclsph.hpp:749 */
  }
  F W6(F r) const {
    if (!(r <= h)) return F(0);
    const F t = std::fma(-r, r, hh);
    return poly6 * ((t * t) * t);
  }
  /* gradient scale s with gradW = d * s */
  F gradScale(F r) const {
    if (!(r >= F(1e-8) && r <= h)) return F(0);
    const F q = h - r;
    return spiky * ((q * q) / r);
  }
};

template <typename F> inline F dist2(F dx, F dy, F dz) {
  return std::fma(dz, dz, std::fma(dy, dy, dx * dx));
}

/* Unpack ARGB (clutils.hpp:224-229) into 4 channel floats, s0 = A. */
template <typename F> inline void unpackARGB(uint32_t c, F out[4]) {
  out[0] = F((c >> 24) & 0xFF);
  out[1] = F((c >> 16) & 0xFF);
  out[2] = F((c >> 8) & 0xFF);
  out[3] = F(c & 0xFF);
}
/* packARGB (clutils.hpp:205-210) of a uchar4. */
inline uint32_t packARGB(const uint8_t c[4]) {
  return (uint32_t(c[0]) << 24) | (uint32_t(c[1]) << 16) | (uint32_t(c[2]) << 8) |
         uint32_t(c[3]);
}
/* convert_uchar4 of a float in [0,255]: round toward zero. */
template <typename F> inline uint8_t toU8(F v) {
  if (!(v > F(0))) return 0;
  if (v >= F(255)) return 255;
  return static_cast<uint8_t>(v);
}

template <typename F> class Pbf {
 public:
  explicit Pbf(const lf_pbf_params& p) : c(p), params(p) {}

  Consts<F> c;
  lf_pbf_params params;
  size_t n = 0;
  /* caller-order state (world units) */
  std::vector<F> x, v, mass;
  std::vector<int32_t> type;
  std::vector<uint32_t> colour;
  std::vector<lf_well> wells;

  /* Mesh colliders (oclsph_collision.h): per-triangle values of
   * collideTriangle2 / inTrig folded once, dot3 = fma(u.z,v.z, fma(u.y,v.y,
   * u.x*v.x)), the unit normal nn = n / |n| of n = (b-a) x (c-a). */
  struct ColliderTri {
    F a[3], nn[3], v0[3], v1[3], dn, dot00, dot01, dot11, invDenom;
  };
  std::vector<ColliderTri> mesh;
  static F dot3(const F* u, const F* v) { return std::fma(u[2], v[2], std::fma(u[1], v[1], u[0] * v[0])); }
  void setColliders(size_t nt, const lf_triangle* tris) {
    mesh.resize(nt);
    for (size_t i = 0; i < nt; ++i) {
      ColliderTri& T = mesh[i];
      F e1[3], e2[3], nv[3];
      for (int k = 0; k < 3; ++k) {
        T.a[k] = F(tris[i].v0[k]);
        e1[k] = F(tris[i].v1[k]) - T.a[k];
        e2[k] = F(tris[i].v2[k]) - T.a[k];
      }
      nv[0] = e1[1] * e2[2] - e1[2] * e2[1];
      nv[1] = e1[2] * e2[0] - e1[0] * e2[2];
      nv[2] = e1[0] * e2[1] - e1[1] * e2[0];
      const F len = std::sqrt(dot3(nv, nv));
      for (int k = 0; k < 3; ++k) {
        T.nn[k] = nv[k] / len;
        T.v0[k] = e2[k];
        T.v1[k] = e1[k];
      }
      T.dn = dot3(T.nn, T.a);
      T.dot00 = dot3(T.v0, T.v0);
      T.dot01 = dot3(T.v0, T.v1);
      T.dot11 = dot3(T.v1, T.v1);
      T.invDenom = F(1) / (T.dot00 * T.dot11 - T.dot01 * T.dot01);
    }
  }
  /* collideTriangle2 (oclsph_collision.h:25-58) with inTrig (:7-22). */
  bool collide(F p[3], F v[3]) const {
    for (const ColliderTri& T : mesh) {
      const F t = T.dn - dot3(T.nn, p);
      if (!(F(-40) <= t && t <= F(40))) continue;
      F p0[3], v2[3];
      for (int k = 0; k < 3; ++k) {
        p0[k] = std::fma(T.nn[k], t, p[k]);
        v2[k] = p0[k] - T.a[k];
      }
      const F dot02 = dot3(T.v0, v2), dot12 = dot3(T.v1, v2);
      const F u = (T.dot11 * dot02 - T.dot01 * dot12) * T.invDenom;
      const F w = (T.dot00 * dot12 - T.dot01 * dot02) * T.invDenom;
      if (!(u >= F(0) && w >= F(0) && u + w < F(1))) continue;
      const F sv = F(2) * dot3(v, T.nn);
      for (int k = 0; k < 3; ++k) {
        p[k] = p0[k];
        v[k] = std::fma(-T.nn[k], sv, v[k]);
      }
      return true;
    }
    return false;
  }
  /* sorted state of the current step: keys = sort keys min(key, G), cellKeys
   * = the exact Morton keys the 27-cell walks decode */
  std::vector<uint32_t> keyIn, keys, cellKeys, order, table;
  std::vector<F> ps, psNext, xs, vs, ms, lambda, rho;
  std::vector<int32_t> ts;
  std::vector<uint32_t> cs, diffused;

  void load(size_t count, const F* pos, const F* vel, const F* m, const int32_t* t,
            const uint32_t* col) {
    n = count;
    x.assign(pos, pos + 3 * n);
    v.assign(vel, vel + 3 * n);
    mass.assign(m, m + n);
    type.assign(t, t + n);
    colour.assign(col, col + n);
  }

  /* advectAndCopy (oclsph.hpp:289-325): returns p* and the post-gravity v. */
  void predict(size_t i, F pstar[3]) const {
    const F* xi = &x[3 * i];
    if (type[i] == LF_OBSTACLE) {
      for (int k = 0; k < 3; ++k) pstar[k] = xi[k] / c.scale;
      return;
    }
    F force[3];
    for (int k = 0; k < 3; ++k) force[k] = mass[i] * F(params.gravity[k]);
    for (const lf_well& w : wells) { /* oclsph.hpp:307-317 */
      F d[3];
      for (int k = 0; k < 3; ++k) d[k] = F(w.centre[k]) - xi[k];
      const F dd = dist2(d[0], d[1], d[2]);
      const F dist = std::sqrt(dd);
      if (dist < F(75)) {
        const F inv = F(1) / dist;
        for (int k = 0; k < 3; ++k) {
          F f = ((d[k] * inv) * F(w.force)) * mass[i];
          f = f / (dist * dist);
          f = std::min(std::max(f, F(-10)), F(10));
          force[k] += f;
        }
      }
    }
    for (int k = 0; k < 3; ++k) {
      const F vn = std::fma(force[k], c.dt, v[3 * i + k]);
      pstar[k] = std::fma(vn, c.dt, xi[k] / c.scale);
    }
  }

  uint32_t keyOf(const F p[3]) const {
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) q[k] = cellCoord<F>((p[k] - c.minExt[k]) / c.h);
    return morton(q[0], q[1], q[2]);
  }

  /* Host pre-pass: predict, key, stable sort, reorder, dense cell table
   * (oclsph.hpp:717-767). */
  void gridBuild() {
    keyIn.resize(n);
    std::vector<F> pIn(3 * n);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < long(n); ++i) {
      predict(size_t(i), &pIn[3 * i]);
      keyIn[i] = keyOf(&pIn[3 * i]);
    }
    order.resize(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
      return std::min(keyIn[a], c.G) < std::min(keyIn[b], c.G);
    });
    keys.resize(n);
    cellKeys.resize(n);
    ps.resize(4 * n);
    xs.resize(3 * n);
    vs.resize(3 * n);
    ms.resize(n);
    ts.resize(n);
    cs.resize(n);
#pragma omp parallel for schedule(static)
    for (long s = 0; s < long(n); ++s) {
      const uint32_t i = order[s];
      keys[s] = std::min(keyIn[i], c.G);
      cellKeys[s] = keyIn[i];
      for (int k = 0; k < 3; ++k) {
        ps[4 * s + k] = pIn[3 * i + k];
        xs[3 * s + k] = x[3 * i + k];
        vs[3 * s + k] = v[3 * i + k]; /* pre-gravity v: oclsph.hpp:300,831 */
      }
      ps[4 * s + 3] = mass[i];
      ms[s] = mass[i];
      ts[s] = type[i];
      cs[s] = colour[i];
    }
    /* oclsph.hpp:758-765 */
    table.assign(c.G, 0u);
    uint32_t gi = 0;
    for (uint32_t cell = 0; cell < c.G; ++cell) {
      table[cell] = gi;
      while (gi != n && keys[gi] == cell) ++gi;
    }
    lambda.assign(n, F(0));
    rho.assign(n, F(0));
    diffused = cs;
  }

  /* FOR_EACH_NEIGHBOUR__ (oclsph_kernel.h:150-192). */
  template <typename Fn> void forEachCandidate(uint32_t key, Fn&& fn) const {
    uint32_t cx, cy, cz;
    demorton(key, cx, cy, cz);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const uint32_t code = morton(uint32_t(int(cx) + dx) & 1023u,
                                       uint32_t(int(cy) + dy) & 1023u,
                                       uint32_t(int(cz) + dz) & 1023u);
          if (size_t(code) + 1 >= c.G) continue; /* end = start */
          for (uint32_t b = table[code]; b < table[code + 1]; ++b) fn(b);
        }
  }

  /* sph_lambda (oclsph_kernel.h:305-332). */
  void lambdaPass() {
    const bool kahan = (c.flags & LF_FLAG_KAHAN) != 0;
#pragma omp parallel for schedule(dynamic, 256)
    for (long al = 0; al < long(n); ++al) {
      const size_t a = size_t(al);
      if (ts[a] == LF_OBSTACLE) {
        lambda[a] = F(0);
        rho[a] = F(0);
        continue;
      }
      const F* pa = &ps[4 * a];
      const F ma = ps[4 * a + 3];
      F g[3] = {0, 0, 0}, gc[3] = {0, 0, 0}, r = 0, rc = 0;
      forEachCandidate(cellKeys[a], [&](uint32_t b) {
        const F* pb = &ps[4 * b];
        const F dx = pa[0] - pb[0], dy = pa[1] - pb[1], dz = pa[2] - pb[2];
        const F d = std::sqrt(dist2(dx, dy, dz));
        const F w = c.W6(d);
        const F s = c.gradScale(d);
        const F gv[3] = {dx * s, dy * s, dz * s};
        if (!kahan) {
          for (int k = 0; k < 3; ++k) g[k] = std::fma(gv[k], c.rhoRecip, g[k]);
          r = std::fma(ma, w, r);
        } else { /* Kahan-Babuska (Neumaier) compensated sums */
          for (int k = 0; k < 3; ++k) {
            const F term = gv[k] * c.rhoRecip;
            const F t = g[k] + term;
            gc[k] += (std::fabs(g[k]) >= std::fabs(term)) ? (g[k] - t) + term
                                                          : (term - t) + g[k];
            g[k] = t;
          }
          const F term = ma * w;
          const F t = r + term;
          rc += (std::fabs(r) >= std::fabs(term)) ? (r - t) + term : (term - t) + r;
          r = t;
        }
      });
      if (kahan) {
        for (int k = 0; k < 3; ++k) g[k] += gc[k];
        r += rc;
      }
      const F norm2 = dist2(g[0], g[1], g[2]);
      const F C1 = r / c.rho0 - F(1);
      lambda[a] = -C1 / (norm2 + c.cfm);
      rho[a] = r;
    }
  }

  /* sph_delta (oclsph_kernel.h:335-384) with the mesh colliders, Jacobi. */
  void deltaPass() {
    psNext.resize(4 * n);
#pragma omp parallel for schedule(dynamic, 256)
    for (long al = 0; al < long(n); ++al) {
      const size_t a = size_t(al);
      const F* pa = &ps[4 * a];
      if (ts[a] == LF_OBSTACLE) {
        for (int k = 0; k < 4; ++k) psNext[4 * a + k] = pa[k];
        continue;
      }
      const F la = lambda[a];
      F acc[3] = {0, 0, 0};
      forEachCandidate(cellKeys[a], [&](uint32_t b) {
        const F* pb = &ps[4 * b];
        const F dx = pa[0] - pb[0], dy = pa[1] - pb[1], dz = pa[2] - pb[2];
        const F d = std::sqrt(dist2(dx, dy, dz));
        const F q = c.W6(d) * c.wdqRecip;
        const F corr = -c.corrK * ((q * q) * (q * q));
        const F factor = ((la + lambda[b]) + corr) * c.rhoRecip;
        const F s = c.gradScale(d);
        acc[0] = std::fma(dx * s, factor, acc[0]);
        acc[1] = std::fma(dy * s, factor, acc[1]);
        acc[2] = std::fma(dz * s, factor, acc[2]);
      });
      F pos[3];
      for (int k = 0; k < 3; ++k) pos[k] = (pa[k] + acc[k]) * c.scale;
      if (!mesh.empty()) collide(pos, &vs[3 * a]); /* :364-368, velocity[a] = resp.velocity */
      for (int k = 0; k < 3; ++k) {
        const F q = std::min(c.maxB[k], std::max(c.minB[k], pos[k])); /* :371 */
        psNext[4 * a + k] = q / c.scale;
      }
      psNext[4 * a + 3] = pa[3];
    }
    ps.swap(psNext);
  }

  /* sph_finalise (oclsph_kernel.h:404-416). */
  void finalise() {
#pragma omp parallel for schedule(static)
    for (long al = 0; al < long(n); ++al) {
      const size_t a = size_t(al);
      if (ts[a] == LF_OBSTACLE) continue;
      for (int k = 0; k < 3; ++k) {
        const F p = ps[4 * a + k];
        const F dX = p - xs[3 * a + k] / c.scale;
        xs[3 * a + k] = p * c.scale;
        vs[3 * a + k] = std::fma(dX, c.invDt, vs[3 * a + k]) * c.vd;
      }
    }
  }

  /* XSPH viscosity (absent in the reference; Macklin & Mueller 2013 eq. 17
   * with the SPH volume m_b / rho0 on each term, so c is a dimensionless
   * smoothing fraction): v_a += (c / rho0) sum_b m_b (v_b - v_a) W6(|p_a - p_b|)
   * over b within h (obstacles with zero weight), Jacobi. */
  void xsphPass() {
    if (c.xsph == F(0)) return;
    std::vector<F> out(vs);
#pragma omp parallel for schedule(dynamic, 256)
    for (long al = 0; al < long(n); ++al) {
      const size_t a = size_t(al);
      if (ts[a] == LF_OBSTACLE) continue;
      const F* pa = &ps[4 * a];
      const F* va = &vs[3 * a];
      F acc[3] = {0, 0, 0};
      forEachCandidate(cellKeys[a], [&](uint32_t b) {
        const F* pb = &ps[4 * b];
        const F dx = pa[0] - pb[0], dy = pa[1] - pb[1], dz = pa[2] - pb[2];
        const F d2 = dist2(dx, dy, dz);
        if (!(std::sqrt(d2) <= c.h)) return;
        // W6 on d2; obstacles carry zero weight (their fluid mass is 0)
        const F t = c.hh - d2;
        const F mw = (ts[b] == LF_OBSTACLE ? F(0) : ms[b]) * (c.poly6 * ((t * t) * t));
        for (int k = 0; k < 3; ++k) acc[k] = std::fma(mw, vs[3 * b + k] - va[k], acc[k]);
      });
      for (int k = 0; k < 3; ++k) out[3 * a + k] = std::fma(c.xsphRho, acc[k], va[k]);
    }
    vs.swap(out);
  }

  /* sph_diffuse (oclsph_kernel.h:281-303). */
  void diffusePass() {
    diffused = cs;
    const F a = c.dt / F(750);
#pragma omp parallel for schedule(dynamic, 256)
    for (long al = 0; al < long(n); ++al) {
      const size_t s = size_t(al);
      if (ts[s] == LF_OBSTACLE) continue;
      F mix[4] = {0, 0, 0, 0};
      int N = 0;
      forEachCandidate(cellKeys[s], [&](uint32_t b) {
        if (ts[b] == LF_OBSTACLE) return;
        F cb[4];
        unpackARGB(cs[b], cb);
        for (int k = 0; k < 4; ++k) mix[k] += cb[k];
        ++N;
      });
      F ca[4];
      unpackARGB(cs[s], ca);
      uint8_t out[4];
      for (int k = 0; k < 4; ++k) {
        const F y = (mix[k] / F(N)) * F(1.33);
        F o = std::fma(y - ca[k], a, ca[k]);
        o = std::fmin(std::fmax(o, F(8)), F(255));
        out[k] = toU8(o);
      }
      diffused[s] = packARGB(out);
    }
  }

  /* One whole step (oclsph.hpp:717-767 then runSphKernel :563-641, plus XSPH),
   * leaving the caller-order state = this step's sorted state, as the
   * write-back of oclsph.hpp:874-892 leaves xs. */
  void step() {
    gridBuild();
    if (c.flags & LF_FLAG_DIFFUSE) diffusePass();
    for (uint32_t it = 0; it < c.iterations; ++it) {
      lambdaPass();
      deltaPass();
    }
    finalise();
    xsphPass();
    commit();
  }

  void commit() {
    for (size_t s = 0; s < n; ++s) {
      for (int k = 0; k < 3; ++k) {
        x[3 * s + k] = xs[3 * s + k];
        v[3 * s + k] = vs[3 * s + k];
      }
      mass[s] = ms[s];
      type[s] = ts[s];
      colour[s] = (ts[s] == LF_OBSTACLE) ? cs[s] : diffused[s];
    }
  }

  /* ---- marching cubes ---------------------------------------------------- */
  /* Reference lattice dims: floor(extent * resolution) + 1 (oclsph.hpp:817-819). */
  void refLatticeDims(F res, uint32_t S[3]) const {
    for (int k = 0; k < 3; ++k)
      S[k] = uint32_t(std::floor(F(c.extent[k]) * res)) + 1u;
  }

  /* mc_lattice (oclsph_kernel.h:421-530) on the final positions xs (world)
   * and diffused colours. Undefined reference cases written as zeros: the
   * skipped last z-curve cell (:455-458) and colour / N with N = 0 (:529). */
  void mcReferenceField(const lf_mc_params& m, std::vector<F>& field,
                        std::vector<F>& normal, std::vector<uint32_t>& col) const {
    const F res = F(m.resolution);
    uint32_t S[3];
    refLatticeDims(res, S);
    const size_t total = size_t(S[0]) * S[1] * S[2];
    field.assign(total, F(0));
    normal.assign(4 * total, F(0));
    col.assign(total, 0u);
    const F step = c.h / res;
    const F thr = c.h * c.scale * F(1);
    const F infl = F(m.particle_influence), psize = F(m.particle_size);
    const F k = (-infl) * psize;
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (long xl = 0; xl < long(S[0]); ++xl)
      for (long yl = 0; yl < long(S[1]); ++yl)
        for (uint32_t z = 0; z < S[2]; ++z) {
          const uint32_t xi = uint32_t(xl), yi = uint32_t(yl);
          const F pos[3] = {F(xi), F(yi), F(z)};
          F a[3];
          for (int q = 0; q < 3; ++q) a[q] = std::fma(pos[q], step, c.minExt[q]) * c.scale;
          const uint32_t cxx = cellCoord<F>(pos[0] / res), cyy = cellCoord<F>(pos[1] / res),
                         czz = cellCoord<F>(pos[2] / res);
          if (cxx == c.extent[0] && cyy == c.extent[1] && czz == c.extent[2]) continue;
          auto cl = [](int vv, uint32_t e) {
            return uint32_t(std::min(std::max(vv, 0), int(e) - 1));
          };
          const uint32_t xs3[3] = {cl(int(cxx) - 1, c.extent[0]), cxx, cl(int(cxx) + 1, c.extent[0])};
          const uint32_t ys3[3] = {cl(int(cyy) - 1, c.extent[1]), cyy, cl(int(cyy) + 1, c.extent[1])};
          const uint32_t zs3[3] = {cl(int(czz) - 1, c.extent[2]), czz, cl(int(czz) + 1, c.extent[2])};
          F val = 0, nrm[3] = {0, 0, 0}, cc[4] = {0, 0, 0, 0};
          uint32_t N = 0;
          for (int iz = 0; iz < 3; ++iz)
            for (int iy = 0; iy < 3; ++iy)
              for (int ix = 0; ix < 3; ++ix) {
                const uint32_t code = morton(xs3[ix], ys3[iy], zs3[iz]);
                if (size_t(code) + 1 >= c.G) continue;
                for (uint32_t b = table[code]; b < table[code + 1]; ++b) {
                  if (ts[b] == LF_OBSTACLE) continue;
                  const F l[3] = {xs[3 * b] - a[0], xs[3 * b + 1] - a[1], xs[3 * b + 2] - a[2]};
                  const F len = std::sqrt(dist2(l[0], l[1], l[2]));
                  if (!(len < thr)) continue;
                  const F den = (infl == F(0.5)) ? std::sqrt(len) : std::pow(len, infl);
                  for (int q = 0; q < 3; ++q) nrm[q] = std::fma(k, l[q] / den, nrm[q]);
                  val += psize / den;
                  F cb[4];
                  unpackARGB(diffused[b], cb);
                  for (int q = 0; q < 4; ++q) cc[q] += cb[q];
                  ++N;
                }
              }
          const F nn = dist2(nrm[0], nrm[1], nrm[2]);
          if (nn > F(0)) {
            const F inv = F(1) / std::sqrt(nn);
            for (int q = 0; q < 3; ++q) nrm[q] *= inv;
          }
          const size_t idx = (size_t(xi) * S[1] + yi) * S[2] + z; /* curves.h:17-20 */
          field[idx] = val;
          for (int q = 0; q < 3; ++q) normal[4 * idx + q] = nrm[q];
          uint8_t o[4] = {0, 0, 0, 0};
          if (N > 0)
            for (int q = 0; q < 4; ++q) o[q] = toU8(cc[q] / F(N));
          col[idx] = packARGB(o);
        }
  }

  /* SPH density field rho(x) = sum over fluid b of m_b poly6 (h^2 - d^2)^3 for
   * |x - p*_b| < h, evaluated as r = fma((m_b poly6) t, t t, r) with
   * t = h^2 - dx^2 - dy^2 - dz^2 taken as one fused multiply-add chain from
   * h^2 (x first), on the lattice
   * x_ijk = origin + (i,j,k) * step (solver units). Candidates: the 27 cells
   * around the sample's own cell, as the PBF passes; terms added in order. */
  void mcDensityField(const uint32_t S[3], const F origin[3], const F step[3],
                      std::vector<F>& field) const {
    const size_t total = size_t(S[0]) * S[1] * S[2];
    field.assign(total, F(0));
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (long xl = 0; xl < long(S[0]); ++xl)
      for (long yl = 0; yl < long(S[1]); ++yl)
        for (uint32_t z = 0; z < S[2]; ++z) {
          const F p[3] = {std::fma(F(xl), step[0], origin[0]),
                          std::fma(F(yl), step[1], origin[1]),
                          std::fma(F(z), step[2], origin[2])};
          F r = 0;
          forEachCandidate(keyOf(p), [&](uint32_t b) {
            if (ts[b] == LF_OBSTACLE) return;
            const F* pb = &ps[4 * b];
            const F dx = p[0] - pb[0], dy = p[1] - pb[1], dz = p[2] - pb[2];
            const F t = std::fma(-dz, dz, std::fma(-dy, dy, std::fma(-dx, dx, c.hh)));
            if (!(t > F(0))) return;
            r = std::fma((pb[3] * c.poly6) * t, t * t, r);
          });
          field[(size_t(xl) * S[1] + size_t(yl)) * S[2] + z] = r;
        }
  }
};

/* Density-field lattice spanning the box: origin = minBound/scale,
 * step = ((maxBound - minBound)/scale) / (S - 1) per axis. */
template <typename F>
inline void densityLattice(const lf_pbf_params& p, const uint32_t S[3], F origin[3],
                           F step[3]) {
  for (int k = 0; k < 3; ++k) {
    origin[k] = F(p.min_bound[k]) / F(p.scale);
    const F span = (F(p.max_bound[k]) - F(p.min_bound[k])) / F(p.scale);
    step[k] = span / F(S[k] > 1 ? S[k] - 1 : 1);
  }
}

/* mc_size (oclsph_kernel.h:545-596): case index per cube, x-major cube index
 * (to3d, curves.h:22-27). */
template <typename F>
inline void mcCases(const uint32_t S[3], const F* field, F iso, std::vector<uint8_t>& ci) {
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const uint32_t M[3] = {S[0] - 1, S[1] - 1, S[2] - 1};
  ci.assign(size_t(M[0]) * M[1] * M[2], 0);
#pragma omp parallel for collapse(2) schedule(static)
  for (long x = 0; x < long(M[0]); ++x)
    for (long y = 0; y < long(M[1]); ++y)
      for (uint32_t z = 0; z < M[2]; ++z) {
        uint32_t c = 0;
        for (int i = 0; i < 8; ++i) {
          const size_t idx = (size_t(x + off[i][0]) * S[1] + size_t(y + off[i][1])) * S[2] +
                             (z + off[i][2]);
          if (field[idx] < iso) c |= 1u << i;
        }
        ci[(size_t(x) * M[1] + size_t(y)) * M[2] + z] = uint8_t(c);
      }
}

inline uint32_t mcTrianglesOfCase(uint32_t ci) {
  return mctable::EdgeTable[ci] == 0 ? 0u : mctable::NumVertsTable[ci] / 3u;
}

/* mc_eval (oclsph_kernel.h:599-735) in cube-index order, triangle order of
 * TriTable within a cube. normal/col may be null (positions only). */
template <typename F>
inline void mcEmit(const uint32_t S[3], const F* field, const F* normal, const uint32_t* col,
                   F iso, const F origin[3], const F step[3], F scale,
                   std::vector<lf_mesh_triangle>& out) {
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  std::vector<uint8_t> ci;
  mcCases(S, field, iso, ci);
  out.clear();
  const uint32_t M[3] = {S[0] - 1, S[1] - 1, S[2] - 1};
  for (size_t cube = 0; cube < ci.size(); ++cube) {
    const uint32_t c = ci[cube];
    if (mctable::NumVertsTable[c] == 0) continue;
    const uint32_t x = uint32_t(cube / (size_t(M[1]) * M[2]));
    const uint32_t y = uint32_t((cube - size_t(x) * M[1] * M[2]) / M[2]);
    const uint32_t z = uint32_t(cube - size_t(x) * M[1] * M[2] - size_t(y) * M[2]);
    F val[8], pos[8][3], nrm[8][3];
    F cc[8][4];
    for (int i = 0; i < 8; ++i) {
      const uint32_t o[3] = {x + off[i][0], y + off[i][1], z + off[i][2]};
      const size_t idx = (size_t(o[0]) * S[1] + o[1]) * S[2] + o[2];
      val[i] = field[idx];
      for (int k = 0; k < 3; ++k) {
        pos[i][k] = std::fma(F(o[k]), step[k], origin[k]) * scale;
        nrm[i][k] = normal ? normal[4 * idx + k] : F(0);
      }
      if (col)
        unpackARGB(col[idx], cc[i]);
      else
        for (int k = 0; k < 4; ++k) cc[i][k] = F(255);
    }
    F ts[12][3], ns[12][3];
    uint8_t cs[12][4];
    const int edges = mctable::EdgeTable[c];
    for (int e = 0; e < 12; ++e) {
      if (!(edges & (1 << e))) continue;
      const int a = mctable::kEdgeCorners[e][0], b = mctable::kEdgeCorners[e][1];
      const F t = (iso - val[a]) / (val[b] - val[a]);
      for (int k = 0; k < 3; ++k) {
        ts[e][k] = std::fma(pos[b][k] - pos[a][k], t, pos[a][k]);
        ns[e][k] = std::fma(nrm[b][k] - nrm[a][k], t, nrm[a][k]);
      }
      for (int k = 0; k < 4; ++k) cs[e][k] = toU8(std::fma(cc[b][k] - cc[a][k], t, cc[a][k]));
    }
    for (int i = 0; mctable::TriTable[c][i] != -1; i += 3) {
      lf_mesh_triangle tri;
      const int e0 = mctable::TriTable[c][i], e1 = mctable::TriTable[c][i + 1],
                e2 = mctable::TriTable[c][i + 2];
      for (int k = 0; k < 3; ++k) {
        tri.v0[k] = float(ts[e0][k]);
        tri.v1[k] = float(ts[e1][k]);
        tri.v2[k] = float(ts[e2][k]);
        tri.n0[k] = float(ns[e0][k]);
        tri.n1[k] = float(ns[e1][k]);
        tri.n2[k] = float(ns[e2][k]);
      }
      tri.c[0] = packARGB(cs[e0]);
      tri.c[1] = packARGB(cs[e1]);
      tri.c[2] = packARGB(cs[e2]);
      out.push_back(tri);
    }
  }
}

}  // namespace oracle

#endif  // LF_PBF_ORACLE_HPP
