/*
 * ORACLE — test infrastructure, not product code.
 * C entry points over oracle::Pbf<float> / oracle::Pbf<double> for ctypes
 * (tests/, the smoke check and bench.py's CPU-baseline leg). Floating arrays
 * cross as double: lossless for the FP32 oracle.
 */
#include <chrono>
#include <cstring>
#include <memory>

#include "pbf_oracle.hpp"

namespace {

struct Handle {
  int fp64;
  std::unique_ptr<oracle::Pbf<float>> f;
  std::unique_ptr<oracle::Pbf<double>> d;
};

template <typename F> void loadT(oracle::Pbf<F>& o, size_t n, const double* pos,
                                 const double* vel, const double* mass,
                                 const int32_t* type, const uint32_t* colour) {
  std::vector<F> p(pos, pos + 3 * n), v(vel, vel + 3 * n), m(mass, mass + n);
  o.load(n, p.data(), v.data(), m.data(), type, colour);
}

template <typename F> void getF(const oracle::Pbf<F>& o, int which, double* dst) {
  const std::vector<F>* src = nullptr;
  switch (which) {
    case LF_ARR_PSTAR: src = &o.ps; break;
    case LF_ARR_LAMBDA: src = &o.lambda; break;
    case LF_ARR_DENSITY: src = &o.rho; break;
    case LF_ARR_POSITION: src = &o.xs; break;
    case LF_ARR_VELOCITY: src = &o.vs; break;
    case 101: src = &o.x; break; /* caller-order position */
    case 102: src = &o.v; break; /* caller-order velocity */
    case 103: src = &o.mass; break;
    default: return;
  }
  for (size_t i = 0; i < src->size(); ++i) dst[i] = double((*src)[i]);
}

template <typename F> void setF(oracle::Pbf<F>& o, int which, const double* src) {
  std::vector<F>* dst = nullptr;
  switch (which) {
    case LF_ARR_PSTAR: dst = &o.ps; break;
    case LF_ARR_LAMBDA: dst = &o.lambda; break;
    case LF_ARR_POSITION: dst = &o.xs; break;
    case LF_ARR_VELOCITY: dst = &o.vs; break;
    default: return;
  }
  for (size_t i = 0; i < dst->size(); ++i) (*dst)[i] = F(src[i]);
}

template <typename F> void getU(const oracle::Pbf<F>& o, int which, uint32_t* dst) {
  const std::vector<uint32_t>* src = nullptr;
  switch (which) {
    case LF_ARR_KEYS: src = &o.keys; break;
    case LF_ARR_ORDER: src = &o.order; break;
    case LF_ARR_TABLE: src = &o.table; break;
    case LF_ARR_DIFFUSED: src = &o.diffused; break;
    case 104: src = &o.keyIn; break; /* caller-order keys */
    case 105: src = &o.colour; break;
    default: return;
  }
  std::memcpy(dst, src->data(), src->size() * sizeof(uint32_t));
}

template <typename F> size_t sizeOf(const oracle::Pbf<F>& o, int which) {
  switch (which) {
    case LF_ARR_PSTAR: return o.ps.size();
    case LF_ARR_POSITION: return o.xs.size();
    case LF_ARR_VELOCITY: return o.vs.size();
    case LF_ARR_TABLE: return o.table.size();
    case 101: return o.x.size();
    case 102: return o.v.size();
    default: return o.n;
  }
}

template <typename F> void stage(oracle::Pbf<F>& o, int s) {
  switch (s) {
    case 0: o.gridBuild(); break;
    case 1: o.lambdaPass(); break;
    case 2: o.deltaPass(); break;
    case 3: o.finalise(); break;
    case 4: o.xsphPass(); break;
    case 5: o.diffusePass(); break;
    case 6: o.commit(); break;
    case 7: o.step(); break;
    default: break;
  }
}

}  // namespace

extern "C" {

void* oc_create(const lf_pbf_params* p, int fp64) {
  auto* h = new Handle();
  h->fp64 = fp64;
  if (fp64)
    h->d.reset(new oracle::Pbf<double>(*p));
  else
    h->f.reset(new oracle::Pbf<float>(*p));
  return h;
}

void oc_destroy(void* hp) { delete static_cast<Handle*>(hp); }

void oc_load(void* hp, size_t n, const double* pos, const double* vel, const double* mass,
             const int32_t* type, const uint32_t* colour) {
  auto* h = static_cast<Handle*>(hp);
  if (h->fp64)
    loadT(*h->d, n, pos, vel, mass, type, colour);
  else
    loadT(*h->f, n, pos, vel, mass, type, colour);
}

void oc_set_wells(void* hp, size_t n, const lf_well* w) {
  auto* h = static_cast<Handle*>(hp);
  if (h->fp64)
    h->d->wells.assign(w, w + n);
  else
    h->f->wells.assign(w, w + n);
}

void oc_set_colliders(void* hp, size_t n, const lf_triangle* t) {
  auto* h = static_cast<Handle*>(hp);
  if (h->fp64)
    h->d->setColliders(n, t);
  else
    h->f->setColliders(n, t);
}

/* collideTriangle2 on one point (p, v in world units; updated in place). */
int oc_collide(void* hp, double* p, double* v) {
  auto* h = static_cast<Handle*>(hp);
  if (h->fp64) return h->d->collide(p, v) ? 1 : 0;
  float pf[3] = {float(p[0]), float(p[1]), float(p[2])}, vf[3] = {float(v[0]), float(v[1]), float(v[2])};
  const int hit = h->f->collide(pf, vf) ? 1 : 0;
  for (int k = 0; k < 3; ++k) {
    p[k] = pf[k];
    v[k] = vf[k];
  }
  return hit;
}

/* stage: 0 grid, 1 lambda, 2 delta, 3 finalise, 4 xsph, 5 diffuse, 6 commit,
 * 7 whole step. Returns the wall time in seconds. */
double oc_stage(void* hp, int s, int repeat) {
  auto* h = static_cast<Handle*>(hp);
  const auto t0 = std::chrono::steady_clock::now();
  for (int r = 0; r < repeat; ++r) {
    if (h->fp64)
      stage(*h->d, s);
    else
      stage(*h->f, s);
  }
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

size_t oc_size(void* hp, int which) {
  auto* h = static_cast<Handle*>(hp);
  return h->fp64 ? sizeOf(*h->d, which) : sizeOf(*h->f, which);
}

void oc_get_f64(void* hp, int which, double* dst) {
  auto* h = static_cast<Handle*>(hp);
  if (h->fp64)
    getF(*h->d, which, dst);
  else
    getF(*h->f, which, dst);
}

void oc_set_f64(void* hp, int which, const double* src) {
  auto* h = static_cast<Handle*>(hp);
  if (h->fp64)
    setF(*h->d, which, src);
  else
    setF(*h->f, which, src);
}

void oc_get_u32(void* hp, int which, uint32_t* dst) {
  auto* h = static_cast<Handle*>(hp);
  if (h->fp64)
    getU(*h->d, which, dst);
  else
    getU(*h->f, which, dst);
}

void oc_get_i32_types(void* hp, int32_t* dst) {
  auto* h = static_cast<Handle*>(hp);
  const auto& t = h->fp64 ? h->d->ts : h->f->ts;
  std::memcpy(dst, t.data(), t.size() * sizeof(int32_t));
}

void oc_grid_info(void* hp, uint32_t* G, uint32_t ext[3], double minExt[3]) {
  auto* h = static_cast<Handle*>(hp);
  for (int k = 0; k < 3; ++k) {
    ext[k] = h->fp64 ? h->d->c.extent[k] : h->f->c.extent[k];
    minExt[k] = h->fp64 ? h->d->c.minExt[k] : double(h->f->c.minExt[k]);
  }
  *G = h->fp64 ? h->d->c.G : h->f->c.G;
}

/* Pair-kernel constants as the oracle folds them: poly6, spiky, wdq, h*h. */
void oc_consts(void* hp, double out[4]) {
  auto* h = static_cast<Handle*>(hp);
  if (h->fp64) {
    out[0] = h->d->c.poly6; out[1] = h->d->c.spiky; out[2] = h->d->c.wdq; out[3] = h->d->c.hh;
  } else {
    out[0] = h->f->c.poly6; out[1] = h->f->c.spiky; out[2] = h->f->c.wdq; out[3] = h->f->c.hh;
  }
}

double oc_w6(void* hp, double r) {
  auto* h = static_cast<Handle*>(hp);
  return h->fp64 ? h->d->c.W6(r) : double(h->f->c.W6(float(r)));
}

double oc_grad_scale(void* hp, double r) {
  auto* h = static_cast<Handle*>(hp);
  return h->fp64 ? h->d->c.gradScale(r) : double(h->f->c.gradScale(float(r)));
}

uint32_t oc_morton(uint32_t x, uint32_t y, uint32_t z) { return oracle::morton(x, y, z); }
void oc_demorton(uint32_t k, uint32_t out[3]) { oracle::demorton(k, out[0], out[1], out[2]); }

/* Reference-field lattice dims for resolution res. */
void oc_mc_ref_dims(void* hp, float res, uint32_t S[3]) {
  auto* h = static_cast<Handle*>(hp);
  if (h->fp64)
    h->d->refLatticeDims(res, S);
  else
    h->f->refLatticeDims(res, S);
}

/* mc_lattice on the current sorted state: field[S], normal[4S], colour[S]. */
void oc_mc_reference(void* hp, const lf_mc_params* m, double* field, double* normal,
                     uint32_t* colour) {
  auto* h = static_cast<Handle*>(hp);
  auto run = [&](auto& o) {
    using F = typename std::remove_reference<decltype(o.c.h)>::type;
    std::vector<F> f, nrm;
    std::vector<uint32_t> col;
    o.mcReferenceField(*m, f, nrm, col);
    for (size_t i = 0; i < f.size(); ++i) field[i] = double(f[i]);
    for (size_t i = 0; i < nrm.size(); ++i) normal[i] = double(nrm[i]);
    std::memcpy(colour, col.data(), col.size() * sizeof(uint32_t));
  };
  if (h->fp64)
    run(*h->d);
  else
    run(*h->f);
}

/* Density field on the S lattice spanning the box. */
void oc_mc_density(void* hp, const uint32_t S[3], double* field) {
  auto* h = static_cast<Handle*>(hp);
  auto run = [&](auto& o) {
    using F = typename std::remove_reference<decltype(o.c.h)>::type;
    F origin[3], step[3];
    oracle::densityLattice<F>(o.params, S, origin, step);
    std::vector<F> f;
    o.mcDensityField(S, origin, step, f);
    for (size_t i = 0; i < f.size(); ++i) field[i] = double(f[i]);
  };
  if (h->fp64)
    run(*h->d);
  else
    run(*h->f);
}

void oc_density_lattice(const lf_pbf_params* p, const uint32_t S[3], double origin[3],
                        double step[3], int fp64) {
  if (fp64) {
    oracle::densityLattice<double>(*p, S, origin, step);
  } else {
    float o[3], s[3];
    oracle::densityLattice<float>(*p, S, o, s);
    for (int k = 0; k < 3; ++k) { origin[k] = o[k]; step[k] = s[k]; }
  }
}

void oc_mc_cases(const uint32_t S[3], const double* field, double iso, int fp64, uint8_t* ci) {
  const size_t total = size_t(S[0]) * S[1] * S[2];
  std::vector<uint8_t> out;
  if (fp64) {
    oracle::mcCases<double>(S, field, iso, out);
  } else {
    std::vector<float> f(field, field + total);
    oracle::mcCases<float>(S, f.data(), float(iso), out);
  }
  std::memcpy(ci, out.data(), out.size());
}

/* mc_eval in cube order. normal (4 per sample) and colour may be null.
 * Returns the triangle count; writes at most cap triangles. */
size_t oc_mc_emit(const uint32_t S[3], const double* field, const double* normal,
                  const uint32_t* colour, double iso, const double origin[3],
                  const double step[3], double scale, int fp64, lf_mesh_triangle* out,
                  size_t cap) {
  const size_t total = size_t(S[0]) * S[1] * S[2];
  std::vector<lf_mesh_triangle> tris;
  if (fp64) {
    oracle::mcEmit<double>(S, field, normal, colour, iso, origin, step, scale, tris);
  } else {
    std::vector<float> f(field, field + total), nv;
    if (normal) nv.assign(normal, normal + 4 * total);
    const float o[3] = {float(origin[0]), float(origin[1]), float(origin[2])};
    const float s[3] = {float(step[0]), float(step[1]), float(step[2])};
    oracle::mcEmit<float>(S, f.data(), normal ? nv.data() : nullptr, colour, float(iso), o, s,
                          float(scale), tris);
  }
  const size_t nOut = std::min(cap, tris.size());
  if (out && nOut) std::memcpy(out, tris.data(), nOut * sizeof(lf_mesh_triangle));
  return tris.size();
}

int oc_mc_table(int which, int ci, int i) {
  switch (which) {
    case 0: return mctable::EdgeTable[ci];
    case 1: return mctable::TriTable[ci][i];
    case 2: return mctable::NumVertsTable[ci];
    default: return 0;
  }
}

}  // extern "C"
