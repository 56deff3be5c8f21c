// C ABI of the gfx950 PBF + marching-cubes library (include/libfluid_amd.h).
// Each entry point validates its arguments, runs on the context's stream and
// reports errors as status codes with text in lf_last_error; nothing throws
// across the boundary.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "lf_internal.hpp"

namespace lf {

int fail(Ctx* c, int code, const std::string& what) {
  if (c) c->err = what;
  return code;
}

int hipFail(Ctx* c, hipError_t e, const char* where) {
  if (c) c->err = std::string(where) + ": " + hipGetErrorString(e);
  return LF_ERR_HIP;
}

static float hostW6(const PbfConsts& c, float r) {
  if (!(r <= c.h)) return 0.0f;
  const float t = std::fma(-r, r, c.hh);
  return c.poly6 * ((t * t) * t);
}

// Folds the reference's kernel constants (oclsph_kernel.h:20-51) and the grid
// geometry of computeBoundAndZindex (oclsph.hpp:334-387) in FP32.
int foldConsts(const lf_pbf_params& p, float skinFrac, PbfConsts& c, std::string& err) {
  if (!(p.h > 0.0f) || !(p.dt > 0.0f) || !(p.scale > 0.0f) || !(p.rest_density > 0.0f)) {
    err = "lf_pbf_params: h, dt, scale and rest_density must be > 0";
    return LF_ERR_ARG;
  }
  const float pi = 3.14159265358979323846f;
  c.h = p.h;
  c.hh = c.h * c.h;
  const float hhh = c.hh * c.h;
  c.poly6 = 315.0f / ((64.0f * pi) * ((hhh * hhh) * hhh));
  c.spiky = -(45.0f / ((pi * hhh) * hhh));
  c.rho0 = p.rest_density;
  c.rhoRecip = 1.0f / c.rho0;
  c.cfm = p.cfm_epsilon;
  c.corrK = p.corr_k;
  c.dt = p.dt;
  c.invDt = 1.0f / c.dt;
  c.scale = p.scale;
  c.vd = p.vd;
  c.xsph = p.xsph_c;
  c.xsphRho = c.xsph / c.rho0;
  {
    const double skin = double(skinFrac) * double(c.h);
    const double r = double(c.h) + skin;
    c.listR2 = float(r * r * (1.0 + 1e-5));
    c.skinSafe = float(skin * (1.0 - 1e-5));
  }
  c.flags = p.flags;
  c.wdq = hostW6(c, p.corr_dq * c.h);
  // rmax2: the largest float x with sqrtf(x) <= h
  float x = c.hh;
  while (std::sqrt(x) > c.h) x = std::nextafter(x, 0.0f);
  while (std::sqrt(std::nextafter(x, INFINITY)) <= c.h) x = std::nextafter(x, INFINITY);
  c.rmax2 = x;
  const float pad = c.h * 2.0f;
  for (int k = 0; k < 3; ++k) {
    c.minB[k] = p.min_bound[k];
    c.maxB[k] = p.max_bound[k];
    c.grav[k] = p.gravity[k];
    c.minExt[k] = (p.min_bound[k] / c.scale) - pad;
    const float maxExt = (p.max_bound[k] / c.scale) + pad;
    const float e = (maxExt - c.minExt[k]) / c.h;
    if (!(e >= 0.0f) || e >= 1024.0f) {
      err = "grid extent must be below 1024 cells per axis (10-bit Morton axis)";
      return LF_ERR_MORTON;
    }
    c.extent[k] = uint32_t(e);
  }
  c.G = mortonEncode(c.extent[0], c.extent[1], c.extent[2]);
  return LF_OK;
}

void timingBegin(Ctx* c, const char* name, PendingTiming& t) {
  t.name = name;
  for (int i = 0; i < 2; ++i) {
    hipEvent_t e;
    if (!c->eventPool.empty()) {
      e = c->eventPool.back();
      c->eventPool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    (i == 0 ? t.start : t.stop) = e;
  }
  if (t.start) (void)hipEventRecord(t.start, c->stream);
}

void timingEnd(Ctx* c, PendingTiming& t) {
  if (t.stop) (void)hipEventRecord(t.stop, c->stream);
  if (t.start && t.stop) c->pending.push_back(t);
}

static void drainTimings(Ctx* c) {
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (PendingTiming& t : c->pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
      KernelStat& s = c->stats[t.name];
      s.count += 1;
      s.ms += ms;
    }
    c->eventPool.push_back(t.start);
    c->eventPool.push_back(t.stop);
  }
  c->pending.clear();
}

static int setParams(Ctx* c, const lf_pbf_params* p) {
  if (!p) return fail(c, LF_ERR_ARG, "null lf_pbf_params");
  PbfConsts k{};
  std::string err;
  const int rc = foldConsts(*p, c->skinFrac, k, err);
  if (rc) return fail(c, rc, err);
  c->consts = k;
  return LF_OK;
}

static int requireGrid(Ctx* c, const char* who) {
  if (!c->gridValid) return fail(c, LF_ERR_STATE, std::string(who) + ": call lf_grid_build first");
  if (c->consts.G != c->gridG)
    return fail(c, LF_ERR_STATE, std::string(who) + ": parameters change the grid; rebuild it");
  return LF_OK;
}

static int pbfStepOnce(Ctx* c) {
  int rc;
  if ((rc = gridBuild(c))) return rc;
  if ((rc = pbfBuildLists(c))) return rc;
  if (c->consts.flags & LF_FLAG_DIFFUSE)
    if ((rc = pbfDiffuse(c))) return rc;
  for (uint32_t it = 0; it < c->params.iterations; ++it) {
    if ((rc = pbfLambda(c))) return rc;
    if ((rc = pbfDelta(c))) return rc;
  }
  if ((rc = pbfFinalise(c))) return rc;
  if ((rc = pbfXsph(c))) return rc;
  c->stepDone = true;
  return LF_OK;
}

}  // namespace lf

using lf::Ctx;
using lf::fail;

struct lf_ctx : lf::Ctx {};

extern "C" {

void lf_pbf_params_default(lf_pbf_params* p, float h) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->h = h;
  p->dt = 1.0f / 120.0f;
  p->scale = 1.0f;
  p->iterations = 4;
  p->max_bound[0] = p->max_bound[1] = p->max_bound[2] = 1.0f;
  p->rest_density = 6378.0f;
  p->cfm_epsilon = 600.0f;
  p->corr_k = 0.0001f;
  p->corr_dq = 0.3f;
  p->vd = 0.49f;
  p->xsph_c = 0.0f;
  p->flags = 0;
}

int lf_ctx_create(int device, lf_ctx** out) {
  if (!out) return LF_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return LF_ERR_NODEVICE;
  if (device < 0 || device >= count) return LF_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return LF_ERR_HIP;
  lf_ctx* c = new (std::nothrow) lf_ctx();
  if (!c) return LF_ERR_ARG;
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return LF_ERR_HIP;
  }
  *out = c;
  return LF_OK;
}

int lf_ctx_destroy(lf_ctx* c) {
  if (!c) return LF_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& t : c->pending) {
    (void)hipEventDestroy(t.start);
    (void)hipEventDestroy(t.stop);
  }
  for (hipEvent_t e : c->eventPool) (void)hipEventDestroy(e);
  for (int i = 0; i < 2; ++i) {
    lf::release(c->X[i]);
    lf::release(c->V[i]);
    lf::release(c->T[i]);
    lf::release(c->ID[i]);
    lf::release(c->sortKey[i]);
    lf::release(c->sortVal[i]);
  }
  lf::release(c->P);
  lf::release(c->PL);
  lf::release(c->nbr);
  lf::release(c->nbrCount);
  lf::release(c->P0);
  lf::release(c->maxDisp);
  lf::release(c->cellDisp);
  lf::release(c->rho);
  lf::release(c->diffused);
  lf::release(c->table);
  lf::release(c->sortHist);
  lf::release(c->wells);
  lf::release(c->field);
  lf::release(c->mcNormal);
  lf::release(c->mcColour);
  lf::release(c->mcCases);
  lf::release(c->mcBlock);
  lf::release(c->mcTotal);
  lf::release(c->triPos);
  lf::release(c->triNrm);
  lf::release(c->triCol);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return LF_OK;
}

const char* lf_last_error(const lf_ctx* c) { return c ? c->err.c_str() : "null context"; }

int lf_sync(lf_ctx* c) {
  if (!c) return LF_ERR_ARG;
  LF_HIP(c, hipStreamSynchronize(c->stream));
  return LF_OK;
}

int lf_state_upload(lf_ctx* c, size_t n, const float* pos, const float* vel, const float* mass,
                    const int32_t* type, const uint32_t* colour) {
  if (!c) return LF_ERR_ARG;
  if (n > 0 && (!pos || !vel || !mass || !type || !colour))
    return fail(c, LF_ERR_ARG, "lf_state_upload: null array");
  if (n >= (size_t(1) << 31)) return fail(c, LF_ERR_ARG, "lf_state_upload: n too large");
  LF_HIP(c, hipSetDevice(c->device));
  int rc;
  const int cur = c->cur;
  if ((rc = lf::ensure(c, c->X[cur], n, "positions"))) return rc;
  if ((rc = lf::ensure(c, c->V[cur], n, "velocities"))) return rc;
  if ((rc = lf::ensure(c, c->T[cur], n, "types"))) return rc;
  if ((rc = lf::ensure(c, c->ID[cur], n, "ids"))) return rc;
  std::vector<float4> X(n), V(n);
  std::vector<uint8_t> T(n);
  std::vector<uint32_t> ID(n);
  for (size_t i = 0; i < n; ++i) {
    if (type[i] != LF_FLUID && type[i] != LF_OBSTACLE)
      return fail(c, LF_ERR_ARG, "lf_state_upload: unknown particle type");
    X[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], mass[i]);
    float cw;
    std::memcpy(&cw, &colour[i], 4);
    V[i] = make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], cw);
    T[i] = uint8_t(type[i]);
    ID[i] = uint32_t(i);
  }
  // synchronous copies: the host vectors die at return
  if (n) {
    LF_HIP(c, hipMemcpyAsync(c->X[cur].p, X.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    LF_HIP(c, hipMemcpyAsync(c->V[cur].p, V.data(), n * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    LF_HIP(c, hipMemcpyAsync(c->T[cur].p, T.data(), n, hipMemcpyHostToDevice, c->stream));
    LF_HIP(c, hipMemcpyAsync(c->ID[cur].p, ID.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    LF_HIP(c, hipStreamSynchronize(c->stream));
  }
  c->n = n;
  c->gridValid = false;
  c->stepDone = false;
  c->mcSampled = c->mcCounted = c->mcEmitted = false;
  return LF_OK;
}

int lf_state_download(lf_ctx* c, size_t n, float* pos, float* vel, float* mass, int32_t* type,
                      uint32_t* colour, uint32_t* order) {
  if (!c) return LF_ERR_ARG;
  if (n != c->n) return fail(c, LF_ERR_ARG, "lf_state_download: n differs from the state size");
  LF_HIP(c, hipSetDevice(c->device));
  const int cur = c->cur;
  std::vector<float4> X(n), V(n);
  std::vector<uint8_t> T(n);
  if (n) {
    LF_HIP(c, hipMemcpyAsync(X.data(), c->X[cur].p, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    LF_HIP(c, hipMemcpyAsync(V.data(), c->V[cur].p, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    LF_HIP(c, hipMemcpyAsync(T.data(), c->T[cur].p, n, hipMemcpyDeviceToHost, c->stream));
    if (order)
      LF_HIP(c, hipMemcpyAsync(order, c->ID[cur].p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    LF_HIP(c, hipStreamSynchronize(c->stream));
  }
  for (size_t i = 0; i < n; ++i) {
    if (pos) {
      pos[3 * i] = X[i].x;
      pos[3 * i + 1] = X[i].y;
      pos[3 * i + 2] = X[i].z;
    }
    if (vel) {
      vel[3 * i] = V[i].x;
      vel[3 * i + 1] = V[i].y;
      vel[3 * i + 2] = V[i].z;
    }
    if (mass) mass[i] = X[i].w;
    if (type) type[i] = int32_t(T[i]);
    if (colour) std::memcpy(&colour[i], &V[i].w, 4);
  }
  return LF_OK;
}

int lf_state_size(const lf_ctx* c, size_t* n) {
  if (!c || !n) return LF_ERR_ARG;
  *n = c->n;
  return LF_OK;
}

int lf_set_wells(lf_ctx* c, size_t n, const lf_well* wells) {
  if (!c) return LF_ERR_ARG;
  if (n > 0 && !wells) return fail(c, LF_ERR_ARG, "lf_set_wells: null wells");
  LF_HIP(c, hipSetDevice(c->device));
  if (n > 0) {
    int rc = lf::ensure(c, c->wells, n, "wells");
    if (rc) return rc;
    LF_HIP(c, hipMemcpyAsync(c->wells.p, wells, n * sizeof(lf_well), hipMemcpyHostToDevice, c->stream));
    LF_HIP(c, hipStreamSynchronize(c->stream));
  }
  c->nWells = int(n);
  return LF_OK;
}

int lf_grid_build(lf_ctx* c, const lf_pbf_params* p) {
  if (!c) return LF_ERR_ARG;
  int rc = lf::setParams(c, p);
  if (rc) return rc;
  c->params = *p;
  LF_HIP(c, hipSetDevice(c->device));
  c->stepDone = false;
  if ((rc = lf::gridBuild(c))) return rc;
  return lf::pbfBuildLists(c);
}

#define LF_STAGE(fn, who)                                  \
  if (!c) return LF_ERR_ARG;                               \
  int rc = lf::setParams(c, p);                            \
  if (rc) return rc;                                       \
  if ((rc = lf::requireGrid(c, who))) return rc;           \
  LF_HIP(c, hipSetDevice(c->device));                      \
  return fn(c);

int lf_pbf_diffuse(lf_ctx* c, const lf_pbf_params* p) { LF_STAGE(lf::pbfDiffuse, "lf_pbf_diffuse") }
int lf_pbf_lambda(lf_ctx* c, const lf_pbf_params* p) { LF_STAGE(lf::pbfLambda, "lf_pbf_lambda") }
int lf_pbf_delta(lf_ctx* c, const lf_pbf_params* p) { LF_STAGE(lf::pbfDelta, "lf_pbf_delta") }
int lf_pbf_finalise(lf_ctx* c, const lf_pbf_params* p) { LF_STAGE(lf::pbfFinalise, "lf_pbf_finalise") }
int lf_pbf_xsph(lf_ctx* c, const lf_pbf_params* p) { LF_STAGE(lf::pbfXsph, "lf_pbf_xsph") }

int lf_pbf_step(lf_ctx* c, const lf_pbf_params* p, uint32_t nsteps) {
  if (!c) return LF_ERR_ARG;
  int rc = lf::setParams(c, p);
  if (rc) return rc;
  c->params = *p;
  LF_HIP(c, hipSetDevice(c->device));
  for (uint32_t s = 0; s < nsteps; ++s)
    if ((rc = lf::pbfStepOnce(c))) return rc;
  return LF_OK;
}

int lf_mc_sample(lf_ctx* c, const lf_pbf_params* p, const lf_mc_params* m) {
  if (!c || !m) return LF_ERR_ARG;
  int rc = lf::setParams(c, p);
  if (rc) return rc;
  LF_HIP(c, hipSetDevice(c->device));
  return lf::mcSample(c, *m);
}

int lf_mc_classify_count(lf_ctx* c, const lf_mc_params* m) {
  if (!c || !m) return LF_ERR_ARG;
  LF_HIP(c, hipSetDevice(c->device));
  return lf::mcClassifyCount(c, *m);
}

int lf_mc_emit(lf_ctx* c, const lf_pbf_params* p, const lf_mc_params* m) {
  if (!c || !m) return LF_ERR_ARG;
  if (p) {  // NULL keeps the scale of the last step or lf_mc_field_upload
    const int rc = lf::setParams(c, p);
    if (rc) return rc;
  }
  LF_HIP(c, hipSetDevice(c->device));
  return lf::mcEmit(c, *m);
}

int lf_mc_count(lf_ctx* c, size_t* ntri) {
  if (!c || !ntri) return LF_ERR_ARG;
  if (!c->mcCounted) return fail(c, LF_ERR_STATE, "lf_mc_count: no counts");
  uint32_t t = 0;
  LF_HIP(c, hipMemcpyAsync(&t, c->mcTotal.p, 4, hipMemcpyDeviceToHost, c->stream));
  LF_HIP(c, hipStreamSynchronize(c->stream));
  *ntri = t;
  if (c->mcEmitted && size_t(t) > c->triCap)
    return fail(c, LF_ERR_CAPACITY, "marching cubes produced more triangles than the capacity");
  return LF_OK;
}

int lf_mc_download(lf_ctx* c, lf_mesh_triangle* out, size_t cap, size_t* ntri) {
  if (!c || !ntri) return LF_ERR_ARG;
  if (!c->mcEmitted) return fail(c, LF_ERR_STATE, "lf_mc_download: nothing emitted");
  size_t total = 0;
  int rc = lf_mc_count(c, &total);
  if (rc && rc != LF_ERR_CAPACITY) return rc;
  *ntri = total;
  const size_t n = std::min(std::min(cap, total), c->triCap);
  if (n == 0 || !out) return rc;
  std::vector<float> pos(n * 9), nrm;
  std::vector<uint32_t> col;
  LF_HIP(c, hipMemcpyAsync(pos.data(), c->triPos.p, n * 9 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  const bool attr = c->triNrm.p != nullptr && c->mcAttr;
  if (attr) {
    nrm.resize(n * 9);
    col.resize(n * 3);
    LF_HIP(c, hipMemcpyAsync(nrm.data(), c->triNrm.p, n * 9 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    LF_HIP(c, hipMemcpyAsync(col.data(), c->triCol.p, n * 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  }
  LF_HIP(c, hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < n; ++i) {
    lf_mesh_triangle& t = out[i];
    std::memcpy(t.v0, &pos[9 * i], 3 * sizeof(float));
    std::memcpy(t.v1, &pos[9 * i + 3], 3 * sizeof(float));
    std::memcpy(t.v2, &pos[9 * i + 6], 3 * sizeof(float));
    if (attr) {
      std::memcpy(t.n0, &nrm[9 * i], 3 * sizeof(float));
      std::memcpy(t.n1, &nrm[9 * i + 3], 3 * sizeof(float));
      std::memcpy(t.n2, &nrm[9 * i + 6], 3 * sizeof(float));
      t.c[0] = col[3 * i];
      t.c[1] = col[3 * i + 1];
      t.c[2] = col[3 * i + 2];
    } else {
      std::memset(t.n0, 0, sizeof(t.n0));
      std::memset(t.n1, 0, sizeof(t.n1));
      std::memset(t.n2, 0, sizeof(t.n2));
      t.c[0] = t.c[1] = t.c[2] = 0xFFFFFFFFu;
    }
  }
  return rc;
}

int lf_mc_field_upload(lf_ctx* c, const uint32_t samples[3], const float origin[3],
                       const float step[3], float scale, const float* field) {
  if (!c || !samples || !origin || !step || !field) return LF_ERR_ARG;
  for (int k = 0; k < 3; ++k)
    if (samples[k] < 2) return fail(c, LF_ERR_ARG, "lf_mc_field_upload: samples must be >= 2");
  if (!(scale > 0.0f)) return fail(c, LF_ERR_ARG, "lf_mc_field_upload: scale must be > 0");
  const size_t total = size_t(samples[0]) * samples[1] * samples[2];
  LF_HIP(c, hipSetDevice(c->device));
  int rc = lf::ensure(c, c->field, total, "mc field");
  if (rc) return rc;
  LF_HIP(c, hipMemcpyAsync(c->field.p, field, total * sizeof(float), hipMemcpyHostToDevice, c->stream));
  LF_HIP(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < 3; ++k) {
    c->S[k] = samples[k];
    c->mcOrigin[k] = origin[k];
    c->mcStep[k] = step[k];
  }
  c->consts.scale = scale;
  c->mcAttr = 0;
  c->mcSampled = true;
  c->mcCounted = c->mcEmitted = false;
  return LF_OK;
}

int lf_mc_set_capacity(lf_ctx* c, size_t max_triangles) {
  if (!c || max_triangles == 0) return LF_ERR_ARG;
  if (max_triangles != c->triCap) {
    lf::release(c->triPos);
    lf::release(c->triNrm);
    lf::release(c->triCol);
    c->triCap = max_triangles;
  }
  return LF_OK;
}

int lf_advance(lf_ctx* c, const lf_pbf_params* p, const lf_mc_params* m, lf_particle* xs,
               size_t n, lf_mesh_triangle* tris, size_t tri_cap, size_t* ntri,
               uint32_t* orderOut) {
  if (!c || !p) return LF_ERR_ARG;
  if (ntri) *ntri = 0;
  if (n == 0) return LF_OK;  // oclsph.hpp:711-715: empty input, empty result
  if (!xs) return fail(c, LF_ERR_ARG, "lf_advance: null particles");
  std::vector<float> pos(3 * n), vel(3 * n), mass(n);
  std::vector<int32_t> type(n);
  std::vector<uint32_t> colour(n), order(n);
  for (size_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      pos[3 * i + k] = xs[i].position[k];
      vel[3 * i + k] = xs[i].velocity[k];
    }
    mass[i] = xs[i].mass;
    type[i] = xs[i].type;
    colour[i] = xs[i].colour;
  }
  int rc = lf_state_upload(c, n, pos.data(), vel.data(), mass.data(), type.data(), colour.data());
  if (rc) return rc;
  lf_pbf_params q = *p;
  q.flags |= LF_FLAG_DIFFUSE;  // runSphKernel always diffuses (oclsph.hpp:603-607)
  if ((rc = lf_pbf_step(c, &q, 1))) return rc;
  size_t count = 0;
  if (m) {
    if ((rc = lf_mc_sample(c, &q, m))) return rc;
    if ((rc = lf_mc_classify_count(c, m))) return rc;
    if ((rc = lf_mc_count(c, &count))) return rc;
    if (count > c->triCap) {
      if ((rc = lf_mc_set_capacity(c, count))) return rc;
    }
    if ((rc = lf_mc_emit(c, &q, m))) return rc;
    size_t got = 0;
    if (tris && tri_cap >= count) {
      if ((rc = lf_mc_download(c, tris, tri_cap, &got))) return rc;
    }
  }
  if (ntri) *ntri = count;
  std::vector<lf_particle> before(xs, xs + n);
  if ((rc = lf_state_download(c, n, pos.data(), vel.data(), mass.data(), type.data(),
                              colour.data(), order.data())))
    return rc;
  if (orderOut) std::memcpy(orderOut, order.data(), n * sizeof(uint32_t));
  // oclsph.hpp:874-892: xs leaves in this step's Morton order
  for (size_t i = 0; i < n; ++i) {
    lf_particle& o = xs[i];
    o = before[order[i]];
    if (o.type == LF_FLUID) {
      for (int k = 0; k < 3; ++k) {
        o.position[k] = pos[3 * i + k];
        o.velocity[k] = vel[3 * i + k];
      }
      o.colour = colour[i];
    }
  }
  return LF_OK;
}

int lf_read(lf_ctx* c, int array, void* dst, size_t bytes) {
  if (!c || (!dst && bytes)) return LF_ERR_ARG;
  LF_HIP(c, hipSetDevice(c->device));
  const void* src = nullptr;
  size_t avail = 0;
  const size_t n = c->n;
  switch (array) {
    case LF_ARR_KEYS: src = c->sortKey[c->keysIdx].p; avail = n * 4; break;
    case LF_ARR_ORDER: src = c->sortVal[c->keysIdx].p; avail = n * 4; break;
    case LF_ARR_TABLE: src = c->table.p; avail = size_t(c->consts.G) * 4; break;
    case LF_ARR_PSTAR: src = c->P.p; avail = n * 16; break;
    case LF_ARR_LAMBDA: src = c->PL.p; avail = n * 16; break;
    case LF_ARR_POSITION: src = c->X[c->cur].p; avail = n * 16; break;
    case LF_ARR_VELOCITY: src = c->V[c->cur].p; avail = n * 16; break;
    case LF_ARR_TYPE: src = c->T[c->cur].p; avail = n; break;
    case LF_ARR_DIFFUSED: src = c->diffused.p; avail = n * 4; break;
    case LF_ARR_DENSITY: src = c->rho.p; avail = n * 4; break;
    case LF_ARR_LISTSTATS: src = c->maxDisp.p; avail = c->maxDisp.p ? 16 : 0; break;
    case LF_ARR_FIELD: src = c->field.p; avail = size_t(c->S[0]) * c->S[1] * c->S[2] * 4; break;
    case LF_ARR_MCNORMAL: src = c->mcNormal.p; avail = size_t(c->S[0]) * c->S[1] * c->S[2] * 16; break;
    case LF_ARR_MCCOLOUR: src = c->mcColour.p; avail = size_t(c->S[0]) * c->S[1] * c->S[2] * 4; break;
    case LF_ARR_CASES:
      src = c->mcCases.p;
      avail = c->keepCases ? size_t(c->S[0] - 1) * (c->S[1] - 1) * (c->S[2] - 1) : 0;
      break;
    default: return fail(c, LF_ERR_ARG, "lf_read: unknown array");
  }
  if (bytes > avail || (bytes && !src)) return fail(c, LF_ERR_ARG, "lf_read: more bytes than the array holds");
  if (bytes) {
    LF_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    LF_HIP(c, hipStreamSynchronize(c->stream));
  }
  return LF_OK;
}

int lf_write(lf_ctx* c, int array, const void* src, size_t bytes) {
  if (!c || (!src && bytes)) return LF_ERR_ARG;
  LF_HIP(c, hipSetDevice(c->device));
  void* dst = nullptr;
  size_t avail = 0;
  const size_t n = c->n;
  switch (array) {
    case LF_ARR_PSTAR: dst = c->P.p; avail = n * 16; break;
    case LF_ARR_LAMBDA: dst = c->PL.p; avail = n * 16; break;
    case LF_ARR_POSITION: dst = c->X[c->cur].p; avail = n * 16; break;
    case LF_ARR_VELOCITY: dst = c->V[c->cur].p; avail = n * 16; break;
    case LF_ARR_DIFFUSED: dst = c->diffused.p; avail = n * 4; break;
    case LF_ARR_FIELD: dst = c->field.p; avail = size_t(c->S[0]) * c->S[1] * c->S[2] * 4; break;
    default: return fail(c, LF_ERR_ARG, "lf_write: array is not writable");
  }
  if (bytes > avail || (bytes && !dst)) return fail(c, LF_ERR_ARG, "lf_write: more bytes than the array holds");
  if (bytes) {
    LF_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    if ((array == LF_ARR_PSTAR || array == LF_ARR_LAMBDA) && c->maxDisp.p) {
      // positions changed outside the delta pass: no neighbour list is usable
      static const uint32_t kInf = 0x7F800000u;
      LF_HIP(c, hipMemcpyAsync(c->maxDisp.p, &kInf, 4, hipMemcpyHostToDevice, c->stream));
      c->listFresh = false;
    }
    LF_HIP(c, hipStreamSynchronize(c->stream));
  }
  return LF_OK;
}

int lf_grid_info(const lf_ctx* c, uint32_t* table_size, uint32_t extent[3], float min_extent[3]) {
  if (!c) return LF_ERR_ARG;
  if (table_size) *table_size = c->consts.G;
  for (int k = 0; k < 3; ++k) {
    if (extent) extent[k] = c->consts.extent[k];
    if (min_extent) min_extent[k] = c->consts.minExt[k];
  }
  return LF_OK;
}

int lf_mc_info(const lf_ctx* c, uint32_t samples[3], float origin[3], float step[3]) {
  if (!c) return LF_ERR_ARG;
  for (int k = 0; k < 3; ++k) {
    if (samples) samples[k] = c->S[k];
    if (origin) origin[k] = c->mcOrigin[k];
    if (step) step[k] = c->mcStep[k];
  }
  return LF_OK;
}

int lf_profile_enable(lf_ctx* c, int on) {
  if (!c) return LF_ERR_ARG;
  c->prof = on != 0;
  return LF_OK;
}

int lf_profile_reset(lf_ctx* c) {
  if (!c) return LF_ERR_ARG;
  lf::drainTimings(c);
  c->stats.clear();
  return LF_OK;
}

int lf_kernel_stats(lf_ctx* c, char* buf, size_t cap) {
  if (!c || !buf || cap == 0) return LF_ERR_ARG;
  lf::drainTimings(c);
  std::string s;
  char line[256];
  for (const auto& kv : c->stats) {
    std::snprintf(line, sizeof(line), "%s %llu %.6f\n", kv.first.c_str(),
                  (unsigned long long)kv.second.count, kv.second.ms);
    s += line;
  }
  std::strncpy(buf, s.c_str(), cap - 1);
  buf[cap - 1] = '\0';
  return s.size() < cap ? LF_OK : LF_ERR_ARG;
}

int lf_copy_probe(lf_ctx* c, size_t bytes, int reps, float* ms) {
  if (!c || !ms || bytes < 16 || reps < 1) return LF_ERR_ARG;
  LF_HIP(c, hipSetDevice(c->device));
  return lf::copyProbe(c, bytes, reps, ms);
}

/* Keep marching-cubes case indices for lf_read(LF_ARR_CASES) (parity tests). */
int lf_mc_keep_cases(lf_ctx* c, int on) {
  if (!c) return LF_ERR_ARG;
  c->keepCases = on != 0;
  return LF_OK;
}

}  // extern "C"
