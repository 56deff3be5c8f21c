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

void dropGraph(Ctx* c) {
  if (c->graph) (void)hipGraphExecDestroy(c->graph);
  c->graph = nullptr;
  c->graphSteps = 0;
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
  c.rnDiv = (c.h >= 1e-4f && c.h <= 1e4f) ? 1u : 0u;
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
    // lists rebuilt by a lambda pass (where particles moved more than the
    // skin) take a wider skin, 3x up to 0.45 h (about 100 entries at rest,
    // within the 128-entry capacity), so later passes rebuild them less often:
    // at 1M over steps 5-25, 1.659 -> 1.607 ms/step (1.5x 1.662, 2x 1.638,
    // 2.5x 1.607, 3.5x 1.612, 4x 1.909: lists overflow)
    const double skinR = std::max(skin, std::min(3.0 * skin, 0.45 * double(c.h)));
    const double rR = double(c.h) + skinR;
    c.listR2R = float(rR * rR * (1.0 + 1e-5));
    c.skinSafeR = float(skinR * (1.0 - 1e-5));
  }
  c.flags = p.flags;
  c.densityTol = p.density_tol > 0.0f ? p.density_tol : 0.0f;
  c.qUnit = c.h / 65536.0f;
  c.qInv = 65536.0 / double(c.h);
  if (p.flags & LF_FLAG_PACKED16) {
    // relative positions carry up to one quantum of error per axis: the list
    // reuse test keeps two quanta of margin (sqrt(3) would do)
    c.skinSafe -= 2.0f * c.qUnit;
    if (!(c.skinSafe > 0.0f)) c.skinSafe = 0.0f;
    c.skinSafeR -= 2.0f * c.qUnit;
    if (!(c.skinSafeR > 0.0f)) c.skinSafeR = 0.0f;
  }
  c.wdq = hostW6(c, p.corr_dq * c.h);
  c.wdqRecip = 1.0f / c.wdq;
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

void timingEvents(Ctx* c, const char* name, PendingTiming& t) {
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
  if (!(t.start && t.stop)) {  // both or none: a launch stamps what it is given
    for (hipEvent_t e : {t.start, t.stop})
      if (e) c->eventPool.push_back(e);
    t.start = t.stop = nullptr;
  }
}

void timingQueue(Ctx* c, PendingTiming& t) {
  if (t.start && t.stop) c->pending.push_back(t);
}

static void drainTimings(Ctx* c) {
  if (c->pending.empty()) return;
  syncStreams(c);
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

// Per-scene constants and the chunk list on the device. A single-scene
// context is a batch of one with keyBase 0 and exact keys (the reference's
// sort order); a batch of several gives scene s the key range [keyBase_s,
// keyBase_s + G_s] with out-of-table keys clamped to G_s. Uploads happen only
// when something changed, never during graph capture.
int prepareScenes(Ctx* c) {
  std::vector<PbfConsts> consts;
  std::vector<uint32_t> begin;
  uint64_t total = 0;
  if (!c->batch) {
    PbfConsts k = c->consts;
    k.keyBase = 0;
    consts.push_back(k);
    begin = {0u, uint32_t(c->n)};
    total = uint64_t(k.G) + 1u;
  } else {
    const size_t S = c->sceneParams.size();
    consts.resize(S);
    for (size_t s = 0; s < S; ++s) {
      std::string err;
      const int rc = foldConsts(c->sceneParams[s], c->skinFrac, consts[s], err);
      if (rc) return fail(c, rc, "scene " + std::to_string(s) + ": " + err);
      consts[s].keyBase = uint32_t(total);
      total += uint64_t(consts[s].G) + 1u;
    }
    begin = c->sceneBegin;
    if (total >= (uint64_t(1) << 30))
      return fail(c, LF_ERR_MORTON, "batch: the scenes' cell tables exceed 2^30 entries in total");
  }
  const bool sameConsts = consts.size() == c->uploadedConsts.size() &&
                          std::memcmp(consts.data(), c->uploadedConsts.data(),
                                      consts.size() * sizeof(PbfConsts)) == 0;
  const bool sameChunks = begin == c->uploadedBegin && c->chunksBuiltFor == c->chunkThreads;
  c->sceneConsts = consts;
  // buckets of the sort: every scene's cells and its bucket of keys >= G; the
  // table a caller reads (LF_ARR_TABLE) is a single scene's G cells, or every
  // bucket of a batch
  c->bucketTotal = uint32_t(total);
  c->tableTotal = c->batch ? uint32_t(total) : consts[0].G;
  c->converge = false;
  for (const PbfConsts& k : consts) c->converge = c->converge || k.densityTol > 0.0f;
  c->pairRn = true;
  for (const PbfConsts& k : consts) c->pairRn = c->pairRn && k.rnDiv != 0u;
  if (!c->batch) c->sceneBegin = begin;
  if (sameConsts && sameChunks && c->dConsts.p && c->chunks.p) return LF_OK;
  if (c->capturing)
    return fail(c, LF_ERR_STATE, "parameters or sizes changed during graph capture");
  int rc;
  dropGraph(c);
  if (!sameConsts) {
    if ((rc = ensure(c, c->dConsts, consts.size(), "scene constants"))) return rc;
    LF_HIP(c, hipMemcpyAsync(c->dConsts.p, consts.data(), consts.size() * sizeof(PbfConsts),
                             hipMemcpyHostToDevice, c->stream));
  }
  std::vector<Chunk> chunks;
  if (!sameChunks) {
    const uint32_t B = c->chunkThreads;
    c->chunkOffHost.assign(begin.size(), 0u);
    for (size_t s = 0; s + 1 < begin.size(); ++s) {
      c->chunkOffHost[s] = uint32_t(chunks.size());
      for (uint32_t a = begin[s]; a < begin[s + 1]; a += B)
        chunks.push_back(
            Chunk{uint32_t(s), a, std::min(a + B, begin[s + 1]), uint32_t(chunks.size())});
    }
    c->chunkOffHost.back() = uint32_t(chunks.size());
    if ((rc = ensure(c, c->chunks, std::max<size_t>(chunks.size(), 1), "chunks"))) return rc;
    if ((rc = ensure(c, c->chunkOff, c->chunkOffHost.size(), "chunk offsets"))) return rc;
    LF_HIP(c, hipMemcpyAsync(c->chunkOff.p, c->chunkOffHost.data(),
                             c->chunkOffHost.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                             c->stream));
    const size_t S = begin.size() - 1;
    if ((rc = ensure(c, c->chunkSums, std::max<size_t>(chunks.size(), 1) * kSumWidth, "chunk sums")))
      return rc;
    if ((rc = ensure(c, c->sceneStats, S * kStatSlots * kSumWidth, "scene statistics"))) return rc;
    if ((rc = ensure(c, c->sceneDone, S, "scene flags"))) return rc;
    if ((rc = ensure(c, c->scenePasses, S, "scene passes"))) return rc;
    LF_HIP(c, hipMemsetAsync(c->sceneStats.p, 0, S * kStatSlots * kSumWidth * sizeof(double), c->stream));
    LF_HIP(c, hipMemsetAsync(c->sceneDone.p, 0, S * sizeof(uint32_t), c->stream));
    LF_HIP(c, hipMemsetAsync(c->scenePasses.p, 0, S * sizeof(uint32_t), c->stream));
    if (!chunks.empty())
      LF_HIP(c, hipMemcpyAsync(c->chunks.p, chunks.data(), chunks.size() * sizeof(Chunk),
                               hipMemcpyHostToDevice, c->stream));
    c->nChunks = uint32_t(chunks.size());
  }
  if (c->persistent && !c->queues.p) {
    if ((rc = ensure(c, c->queues, 2 * kQueueSlots, "work queues"))) return rc;
    LF_HIP(c, hipMemsetAsync(c->queues.p, 0, c->queues.cap * sizeof(uint32_t), c->stream));
  }
  LF_HIP(c, hipStreamSynchronize(c->stream));  // the host vectors above die at return
  c->uploadedConsts = consts;
  c->uploadedBegin = begin;
  c->chunksBuiltFor = c->chunkThreads;
  return LF_OK;
}

Work workOf(Ctx* c, int queueSlot, bool skipDone) {
  Work W{};
  W.halt = nullptr;
  W.ppt = c->chunkPpt;
  W.skip = (skipDone && c->converge) ? c->sceneDone.p : nullptr;
  W.consts = c->dConsts.p;
  W.chunks = c->chunks.p;
  W.nChunks = c->nChunks;
  W.queue = (c->persistent && queueSlot >= 0 && queueSlot < kQueueSlots && c->queues.p)
                ? c->queues.p + 2 * queueSlot
                : nullptr;
  return W;
}

int checkHealth(Ctx* c, bool wait) {
  if (!c->healthHost) return LF_OK;
  if (wait) LF_HIP(c, hipStreamSynchronize(c->stream));
  // the copy k_reorder writes: up to date after a wait, else at most the
  // steps still in flight behind
  const uint32_t bad = c->healthHost[0], full = c->healthHost[1];
  if (bad == 0 && full == 0) return LF_OK;
  return fail(c, LF_ERR_DIVERGED,
              "the state diverged: " + std::to_string(bad) + " particle(s) with a non-finite p*, " +
                  std::to_string(full) + " cell(s) above " + std::to_string(c->guardCell) +
                  " particles (largest " + std::to_string(c->healthHost[2]) +
                  "); the pair passes stop until a new state is uploaded");
}

// Clears the health counters (a new state), after the steps in flight (they
// write the host copy), and the per-cell list displacements (the next grid
// build clears only the cells of the particles it is given).
static int resetHealth(Ctx* c) {
  if (c->health.p) LF_HIP(c, hipMemsetAsync(c->health.p, 0, 4 * sizeof(uint32_t), c->stream));
  if (c->cellDisp.p)
    LF_HIP(c, hipMemsetAsync(c->cellDisp.p, 0, c->cellDisp.cap * sizeof(uint32_t), c->stream));
  LF_HIP(c, hipStreamSynchronize(c->stream));
  if (c->healthHost) std::memset(c->healthHost, 0, 4 * sizeof(uint32_t));
  return LF_OK;
}

static int requireGrid(Ctx* c, const char* who) {
  if (!c->gridValid) return fail(c, LF_ERR_STATE, std::string(who) + ": call lf_grid_build first");
  if (c->consts.G != c->gridG)
    return fail(c, LF_ERR_STATE, std::string(who) + ": parameters change the grid; rebuild it");
  return LF_OK;
}

static int batchOnly(Ctx* c, const char* who) {
  return fail(c, LF_ERR_STATE,
              std::string(who) + ": the context holds a batch (use lf_batch_step / lf_batch_select)");
}

// Sources of oclsph.hpp:678-691, appended at the end of the state: a
// floor(sqrt(rate)) x ceil(sqrt(rate)) lattice per source, spacing h*scale/2,
// offset = centre - (width, 0, depth)/2 * spacing, in the reference's float
// operation order; mass 1, fluid.
static int appendSources(Ctx* c) {
  if (c->sources.empty()) return LF_OK;
  const float spacing = c->params.h * c->params.scale / 2.0f;
  std::vector<float4> X, V;
  for (const lf_source& s : c->sources) {
    const float size = float(std::sqrt(double(s.rate)));  // sqrt of an integer rate is taken in double
    const int width = int(std::floor(size)), depth = int(std::ceil(size));
    const float off[3] = {s.centre[0] - (float(width) / 2.0f) * spacing,
                          s.centre[1] - (0.0f / 2.0f) * spacing,
                          s.centre[2] - (float(depth) / 2.0f) * spacing};
    float cw;
    std::memcpy(&cw, &s.colour, 4);
    for (int x = 0; x < width; ++x)
      for (int z = 0; z < depth; ++z) {
        X.push_back(make_float4(off[0] + float(x) * spacing, off[1] + 0.0f * spacing,
                                off[2] + float(z) * spacing, 1.0f));
        V.push_back(make_float4(s.velocity[0], s.velocity[1], s.velocity[2], cw));
      }
  }
  const size_t k = X.size(), n = c->n;
  if (k == 0) return LF_OK;
  if (n + k >= (size_t(1) << 28)) return fail(c, LF_ERR_ARG, "sources: at most 2^28 - 1 particles per context");
  const int cur = c->cur;
  int rc;
  if ((rc = ensureKeep(c, c->X[cur], n + k, n, "positions"))) return rc;
  if ((rc = ensureKeep(c, c->V[cur], n + k, n, "velocities"))) return rc;
  if ((rc = ensureKeep(c, c->T[cur], n + k, n, "types"))) return rc;
  if ((rc = ensureKeep(c, c->ID[cur], n + k, n, "ids"))) return rc;
  std::vector<uint8_t> T(k, uint8_t(LF_FLUID));
  std::vector<uint32_t> ID(k);
  for (size_t i = 0; i < k; ++i) ID[i] = c->nextId++;
  LF_HIP(c, hipMemcpyAsync(c->X[cur].p + n, X.data(), k * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  LF_HIP(c, hipMemcpyAsync(c->V[cur].p + n, V.data(), k * sizeof(float4), hipMemcpyHostToDevice, c->stream));
  LF_HIP(c, hipMemcpyAsync(c->T[cur].p + n, T.data(), k, hipMemcpyHostToDevice, c->stream));
  LF_HIP(c, hipMemcpyAsync(c->ID[cur].p + n, ID.data(), k * sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
  LF_HIP(c, hipStreamSynchronize(c->stream));
  c->n = n + k;
  c->sceneBegin = {0u, uint32_t(c->n)};
  return LF_OK;
}

static int pbfStepOnce(Ctx* c) {
  int rc;
  if ((rc = appendSources(c))) return rc;
  if ((rc = applyDrains(c))) return rc;
  if ((rc = gridBuild(c))) return rc;
  if (c->nQueries && (rc = answerQueries(c))) return rc;
  if ((rc = pbfBuildLists(c))) return rc;
  if (c->consts.flags & LF_FLAG_DIFFUSE)
    if ((rc = pbfDiffuse(c))) return rc;
  // without the convergence check every pass runs, so the last delta pass can
  // finalise each particle itself
  const uint32_t iters = c->params.iterations;
  const bool fuse = !c->converge && iters > 0;
  for (uint32_t it = 0; it < iters; ++it) {
    if ((rc = pbfLambda(c, it > 0))) return rc;
    if (c->converge && (rc = densityCheck(c, int(it)))) return rc;
    if ((rc = pbfDelta(c, true, fuse && it + 1 == iters))) return rc;
  }
  if (!fuse && (rc = pbfFinalise(c))) return rc;
  if ((rc = pbfXsph(c))) return rc;
  c->stepDone = true;
  return LF_OK;
}

}  // namespace lf

using lf::Ctx;
using lf::fail;


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
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      cus > 0)
    c->persistentBlocks = uint32_t(cus) * 8u;  // 8 workgroups of 256 lanes fill a CU's 32 waves
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return LF_ERR_HIP;
  }
  // the address of lf_mc_mesh_static's block-count array on this device
  if (int rc = lf::mcStaticInit(device, &c->mcScratch)) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return rc;
  }

  *out = c;
  return LF_OK;
}

int lf_ctx_destroy(lf_ctx* c) {
  if (!c) return LF_OK;
  (void)hipSetDevice(c->device);
  lf::syncStreams(c);
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
  lf::release(c->Q);
  lf::release(c->QL);
  lf::release(c->LAM);
  lf::release(c->nbr);
  lf::release(c->nbrCount);
  lf::release(c->P0);
  lf::release(c->listEpoch);
  lf::release(c->listR);
  lf::release(c->maxDisp);
  lf::release(c->cellDisp);
  lf::release(c->Palt);
  lf::release(c->tableAlt);
  lf::release(c->rho);
  lf::release(c->diffused);
  lf::release(c->table);
  lf::release(c->cellCount);
  lf::release(c->bucketTile);
  lf::release(c->bigBuckets);
  lf::release(c->health);
  if (c->healthHost) (void)hipHostFree(c->healthHost);
  lf::release(c->wells);
  lf::release(c->colliders);
  lf::release(c->colliderBounds);
  lf::release(c->colliderCellStart);
  lf::release(c->colliderCellTris);
  lf::release(c->colliderCounts);
  lf::release(c->drains);
  lf::release(c->queryPts);
  lf::release(c->queryOut);
  lf::release(c->compactBlock);
  lf::release(c->wireStage);
  lf::release(c->field);
  lf::release(c->mcNormal);
  lf::release(c->mcColour);
  lf::release(c->mcCases);
  lf::release(c->mcBlock);
  lf::release(c->mcTotal);
  lf::release(c->triPos);
  lf::release(c->triNrm);
  lf::release(c->triCol);
  lf::release(c->dConsts);
  lf::release(c->chunks);
  lf::release(c->queues);
  lf::release(c->localKey);
  lf::release(c->chunkOff);
  lf::release(c->chunkSums);
  lf::release(c->sceneStats);
  lf::release(c->sceneDone);
  lf::release(c->scenePasses);
  lf::dropGraph(c);
  for (hipEvent_t e : {c->evMcIn, c->evMcDone, c->pairGuard.ev, c->pairGuardAlt.ev, c->evPosFinal})
    if (e) (void)hipEventDestroy(e);
  if (c->mcStream) (void)hipStreamDestroy(c->mcStream);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return LF_OK;
}

const char* lf_last_error(const lf_ctx* c) { return c ? c->err.c_str() : "null context"; }

int lf_sync(lf_ctx* c) {
  if (!c) return LF_ERR_ARG;
  LF_HIP(c, hipStreamSynchronize(c->stream));
  if (c->mcStream) LF_HIP(c, hipStreamSynchronize(c->mcStream));
  lf::syncStreams(c);  // clears the pending stream waits
  return lf::checkHealth(c, true);
}

int lf_health(lf_ctx* c, uint32_t out[3]) {
  if (!c || !out) return LF_ERR_ARG;
  LF_HIP(c, hipSetDevice(c->device));
  uint32_t h[4] = {0, 0, 0, 0};
  if (c->health.p) {
    LF_HIP(c, hipMemcpyAsync(h, c->health.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    LF_HIP(c, hipStreamSynchronize(c->stream));
  }
  for (int i = 0; i < 3; ++i) out[i] = h[i];
  return LF_OK;
}

int lf_set_guard(lf_ctx* c, uint32_t max_cell_particles) {
  if (!c) return LF_ERR_ARG;
  LF_HIP(c, hipSetDevice(c->device));
  lf::dropGraph(c);  // a captured graph holds the old limit
  c->guardCell = max_cell_particles;
  return lf::resetHealth(c);
}

int lf_state_upload(lf_ctx* c, size_t n, const float* pos, const float* vel, const float* mass,
                    const int32_t* type, const uint32_t* colour) {
  if (!c) return LF_ERR_ARG;
  c->posEventValid = false;  // P and the table are rewritten outside the event
  if (n > 0 && (!pos || !vel || !mass || !type || !colour))
    return fail(c, LF_ERR_ARG, "lf_state_upload: null array");
  if (n >= (size_t(1) << 28))  // gathers address particles with 32-bit byte offsets
    return fail(c, LF_ERR_ARG, "lf_state_upload: at most 2^28 - 1 particles per context");
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
  c->nextId = uint32_t(n);
  c->queriesAnswered = false;
  c->batch = false;
  c->sceneParams.clear();
  c->sceneBegin = {0u, uint32_t(n)};
  c->mcScene = 0;
  c->gridValid = false;
  c->stepDone = false;
  c->mcSampled = c->mcCounted = c->mcEmitted = false;
  lf::dropGraph(c);
  return lf::resetHealth(c);
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

int lf_set_sources(lf_ctx* c, size_t n, const lf_source* src) {
  if (!c) return LF_ERR_ARG;
  if (n > 0 && !src) return fail(c, LF_ERR_ARG, "lf_set_sources: null sources");
  if (n > 0 && c->batch) return fail(c, LF_ERR_STATE, "lf_set_sources: not for batches");
  c->sources.assign(src, src + n);
  lf::dropGraph(c);
  return LF_OK;
}

int lf_set_drains(lf_ctx* c, size_t n, const lf_drain* d) {
  if (!c) return LF_ERR_ARG;
  if (n > 0 && !d) return fail(c, LF_ERR_ARG, "lf_set_drains: null drains");
  if (n > 0 && c->batch) return fail(c, LF_ERR_STATE, "lf_set_drains: not for batches");
  LF_HIP(c, hipSetDevice(c->device));
  if (n > 0) {
    int rc = lf::ensure(c, c->drains, n, "drains");
    if (rc) return rc;
    LF_HIP(c, hipMemcpyAsync(c->drains.p, d, n * sizeof(lf_drain), hipMemcpyHostToDevice, c->stream));
    LF_HIP(c, hipStreamSynchronize(c->stream));
  }
  c->hostDrains.assign(d, d + n);
  lf::dropGraph(c);
  return LF_OK;
}

int lf_set_queries(lf_ctx* c, size_t n, const float* pts) {
  if (!c) return LF_ERR_ARG;
  if (n > 0 && !pts) return fail(c, LF_ERR_ARG, "lf_set_queries: null points");
  if (n > 0 && c->batch) return fail(c, LF_ERR_STATE, "lf_set_queries: not for batches");
  if (n >= (size_t(1) << 30)) return fail(c, LF_ERR_ARG, "lf_set_queries: too many points");
  LF_HIP(c, hipSetDevice(c->device));
  if (n > 0) {
    int rc = lf::ensure(c, c->queryPts, 3 * n, "query points");
    if (rc) return rc;
    LF_HIP(c, hipMemcpyAsync(c->queryPts.p, pts, 3 * n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    LF_HIP(c, hipStreamSynchronize(c->stream));
  }
  lf::dropGraph(c);
  c->nQueries = uint32_t(n);
  c->queriesAnswered = false;
  return LF_OK;
}

int lf_query_results(lf_ctx* c, size_t n, uint32_t* counts, uint32_t* colours) {
  if (!c) return LF_ERR_ARG;
  if (n != c->nQueries) return fail(c, LF_ERR_ARG, "lf_query_results: n differs from the query count");
  if (!c->queriesAnswered) return fail(c, LF_ERR_STATE, "lf_query_results: no step since lf_set_queries");
  LF_HIP(c, hipSetDevice(c->device));
  std::vector<uint32_t> out(2 * n);
  if (n) {
    LF_HIP(c, hipMemcpyAsync(out.data(), c->queryOut.p, 2 * n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    LF_HIP(c, hipStreamSynchronize(c->stream));
  }
  for (size_t i = 0; i < n; ++i) {
    if (counts) counts[i] = out[2 * i];
    if (colours) colours[i] = out[2 * i + 1];
  }
  return LF_OK;
}

int lf_set_colliders(lf_ctx* c, size_t n, const lf_triangle* tris) {
  if (!c) return LF_ERR_ARG;
  if (n > 0 && !tris) return fail(c, LF_ERR_ARG, "lf_set_colliders: null triangles");
  if (n >= (size_t(1) << 31)) return fail(c, LF_ERR_ARG, "lf_set_colliders: too many triangles");
  LF_HIP(c, hipSetDevice(c->device));
  return lf::setColliders(c, n, tris);
}

int lf_grid_build(lf_ctx* c, const lf_pbf_params* p) {
  if (!c) return LF_ERR_ARG;
  if (c->batch) return lf::batchOnly(c, "lf_grid_build");
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
  if (c->batch) return lf::batchOnly(c, who);              \
  int rc = lf::setParams(c, p);                            \
  if (rc) return rc;                                       \
  if ((rc = lf::requireGrid(c, who))) return rc;           \
  LF_HIP(c, hipSetDevice(c->device));                      \
  if ((rc = lf::prepareScenes(c))) return rc;              \
  return fn(c);

int lf_pbf_diffuse(lf_ctx* c, const lf_pbf_params* p) { LF_STAGE(lf::pbfDiffuse, "lf_pbf_diffuse") }
static int lambdaStage(Ctx* c) { return lf::pbfLambda(c, false); }
static int deltaStage(Ctx* c) { return lf::pbfDelta(c, false); }
int lf_pbf_lambda(lf_ctx* c, const lf_pbf_params* p) { LF_STAGE(lambdaStage, "lf_pbf_lambda") }
int lf_pbf_delta(lf_ctx* c, const lf_pbf_params* p) { LF_STAGE(deltaStage, "lf_pbf_delta") }
int lf_pbf_finalise(lf_ctx* c, const lf_pbf_params* p) { LF_STAGE(lf::pbfFinalise, "lf_pbf_finalise") }
int lf_pbf_xsph(lf_ctx* c, const lf_pbf_params* p) { LF_STAGE(lf::pbfXsph, "lf_pbf_xsph") }

int lf_pbf_step(lf_ctx* c, const lf_pbf_params* p, uint32_t nsteps) {
  if (!c) return LF_ERR_ARG;
  if (c->batch) return lf::batchOnly(c, "lf_pbf_step");
  int rc = lf::setParams(c, p);
  if (rc) return rc;
  c->params = *p;
  LF_HIP(c, hipSetDevice(c->device));
  if ((rc = lf::checkHealth(c, false))) return rc;
  for (uint32_t s = 0; s < nsteps; ++s)
    if ((rc = lf::pbfStepOnce(c))) return rc;
  return LF_OK;
}

int lf_mc_sample(lf_ctx* c, const lf_pbf_params* p, const lf_mc_params* m) {
  if (!c || !m) return LF_ERR_ARG;
  if (p) {  // NULL: the constants of the last step (in a batch: the selected scene's)
    if (c->batch) return fail(c, LF_ERR_ARG, "lf_mc_sample: pass NULL params for a batch scene");
    const int rc = lf::setParams(c, p);
    if (rc) return rc;
  }
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
  if (p) {  // NULL keeps the scale of the last lf_mc_sample or lf_mc_field_upload
    if (!(p->scale > 0.0f)) return fail(c, LF_ERR_ARG, "lf_mc_emit: scale must be > 0");
    c->mcScale = p->scale;
  }
  LF_HIP(c, hipSetDevice(c->device));
  return lf::mcEmit(c, *m);
}

int lf_mc_count(lf_ctx* c, size_t* ntri) {
  if (!c || !ntri) return LF_ERR_ARG;
  if (lf::mcJoin(c)) return lf::fail(c, LF_ERR_HIP, "waiting for the marching-cubes stream");
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
  if (lf::mcJoin(c)) return lf::fail(c, LF_ERR_HIP, "waiting for the marching-cubes stream");
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
  if (lf::mcJoin(c)) return lf::fail(c, LF_ERR_HIP, "waiting for the marching-cubes stream");
  for (int k = 0; k < 3; ++k)
    if (samples[k] < 2) return fail(c, LF_ERR_ARG, "lf_mc_field_upload: samples must be >= 2");
  if (!(scale > 0.0f)) return fail(c, LF_ERR_ARG, "lf_mc_field_upload: scale must be > 0");
  const size_t total = size_t(samples[0]) * samples[1] * samples[2];
  LF_HIP(c, hipSetDevice(c->device));
  int rc = lf::ensure(c, c->field, total, "mc field");
  if (rc) return rc;
  LF_HIP(c, hipMemcpyAsync(c->field.p, field, total * sizeof(float), hipMemcpyHostToDevice, c->stream));
  c->fieldOnMain = true;
  c->mcTileFlags = false;
  LF_HIP(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < 3; ++k) {
    c->S[k] = samples[k];
    c->mcOrigin[k] = origin[k];
    c->mcStep[k] = step[k];
  }
  c->mcScale = scale;
  c->mcAttr = 0;
  c->mcSampled = true;
  c->mcCounted = c->mcEmitted = false;
  return LF_OK;
}

int lf_mc_mesh_static(lf_ctx* c, const uint32_t samples[3], const float origin[3],
                      const float step[3], float scale, float isolevel, const float* d_field,
                      float* d_tri_pos, size_t tri_cap, uint32_t* d_ntri) {
  if (!c || !samples || !origin || !step || !d_field || !d_ntri || (tri_cap && !d_tri_pos))
    return LF_ERR_ARG;
  if (!(scale > 0.0f)) return fail(c, LF_ERR_ARG, "lf_mc_mesh_static: scale must be > 0");
  LF_HIP(c, hipSetDevice(c->device));
  return lf::mcMeshStatic(c, samples, origin, step, scale, isolevel, d_field, d_tri_pos, tri_cap,
                          d_ntri);
}

int lf_mc_set_capacity(lf_ctx* c, size_t max_triangles) {
  if (!c || max_triangles == 0) return LF_ERR_ARG;
  if (lf::mcJoin(c)) return lf::fail(c, LF_ERR_HIP, "waiting for the marching-cubes stream");
  if (max_triangles != c->triCap) {
    // nothing in flight may still write the old buffers, and a captured graph
    // holds their addresses and the old capacity: it goes with them
    LF_HIP(c, hipSetDevice(c->device));
    lf::syncStreams(c);
    lf::dropGraph(c);
    lf::release(c->triPos);
    lf::release(c->triNrm);
    lf::release(c->triCol);
    c->triCap = max_triangles;
    c->mcEmitted = false;  // the triangles went with the buffers; the counts stay valid
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
  if (lf::mcJoin(c)) return lf::fail(c, LF_ERR_HIP, "waiting for the marching-cubes stream");
  LF_HIP(c, hipSetDevice(c->device));
  const void* src = nullptr;
  size_t avail = 0;
  const size_t n = c->n;
  switch (array) {
    case LF_ARR_KEYS: src = c->sortKey[c->keysIdx].p; avail = n * 4; break;
    case LF_ARR_ORDER: src = c->sortVal[c->keysIdx].p; avail = n * 4; break;
    case LF_ARR_TABLE: src = c->table.p; avail = size_t(c->gridTotal) * 4; break;
    case LF_ARR_CELLKEY: src = c->localKey.p; avail = n * 4; break;
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
  // the marching-cubes stream may read what is written; the early position
  // event no longer covers P; a written field is the main stream's
  if (lf::mcJoin(c)) return lf::fail(c, LF_ERR_HIP, "waiting for the marching-cubes stream");
  c->posEventValid = false;
  if (array == LF_ARR_FIELD) {
    c->fieldOnMain = true;
    c->mcTileFlags = false;
  }
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
      // (D and every pass's D_j infinite)
      static const uint32_t kInf[20] = {0x7F800000u, 0, 0, 0, 0x7F800000u, 0x7F800000u,
                                        0x7F800000u, 0x7F800000u, 0x7F800000u, 0x7F800000u,
                                        0x7F800000u, 0x7F800000u, 0x7F800000u, 0x7F800000u,
                                        0x7F800000u, 0x7F800000u, 0x7F800000u, 0x7F800000u,
                                        0x7F800000u, 0x7F800000u};
      LF_HIP(c, hipMemcpyAsync(c->maxDisp.p, &kInf[0], 4, hipMemcpyHostToDevice, c->stream));
      LF_HIP(c, hipMemcpyAsync(c->maxDisp.p + 4, &kInf[4], 16 * 4, hipMemcpyHostToDevice,
                               c->stream));
      c->listFresh = false;
    }
    LF_HIP(c, hipStreamSynchronize(c->stream));
  }
  return LF_OK;
}

int lf_grid_info(const lf_ctx* c, uint32_t* table_size, uint32_t extent[3], float min_extent[3]) {
  if (!c) return LF_ERR_ARG;
  const lf::PbfConsts& k =
      (c->batch && size_t(c->mcScene) < c->sceneConsts.size()) ? c->sceneConsts[c->mcScene] : c->consts;
  if (table_size) *table_size = k.G;
  for (int i = 0; i < 3; ++i) {
    if (extent) extent[i] = k.extent[i];
    if (min_extent) min_extent[i] = k.minExt[i];
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

int lf_profile_only(lf_ctx* c, const char* name) {
  if (!c) return LF_ERR_ARG;
  c->profOnly = name ? name : "";
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

int lf_batch_upload(lf_ctx* c, uint32_t nScenes, const lf_pbf_params* params,
                    const size_t* counts, const float* pos, const float* vel, const float* mass,
                    const int32_t* type, const uint32_t* colour) {
  if (!c) return LF_ERR_ARG;
  c->posEventValid = false;  // P and the table are rewritten outside the event
  if (nScenes == 0 || !params || !counts) return fail(c, LF_ERR_ARG, "lf_batch_upload: no scenes");
  size_t n = 0;
  std::vector<uint32_t> begin(nScenes + 1, 0u);
  for (uint32_t s = 0; s < nScenes; ++s) {
    n += counts[s];
    if (n >= (size_t(1) << 31)) return fail(c, LF_ERR_ARG, "lf_batch_upload: too many particles");
    begin[s + 1] = uint32_t(n);
    const lf_pbf_params& a = params[0];
    const lf_pbf_params& b = params[s];
    if (b.h != a.h || b.iterations != a.iterations || b.flags != a.flags)
      return fail(c, LF_ERR_ARG, "lf_batch_upload: scenes must share h, iterations and flags");
    lf::PbfConsts k{};
    std::string err;
    const int rc = lf::foldConsts(b, c->skinFrac, k, err);
    if (rc) return fail(c, rc, "scene " + std::to_string(s) + ": " + err);
  }
  if (!c->sources.empty() || !c->hostDrains.empty() || c->nQueries)
    return fail(c, LF_ERR_STATE, "lf_batch_upload: sources, drains and queries are single-scene only");
  int rc = lf_state_upload(c, n, pos, vel, mass, type, colour);
  if (rc) return rc;
  c->batch = true;
  c->sceneParams.assign(params, params + nScenes);
  c->sceneBegin = begin;
  c->params = params[0];
  if ((rc = lf::setParams(c, &params[0]))) return rc;
  c->mcScene = 0;
  return LF_OK;
}

int lf_batch_step(lf_ctx* c, uint32_t nsteps) {
  if (!c) return LF_ERR_ARG;
  if (!c->batch) return fail(c, LF_ERR_STATE, "lf_batch_step: no batch (call lf_batch_upload)");
  LF_HIP(c, hipSetDevice(c->device));
  int rc = lf::checkHealth(c, false);
  if (rc) return rc;
  for (uint32_t s = 0; s < nsteps; ++s)
    if ((rc = lf::pbfStepOnce(c))) return rc;
  return LF_OK;
}

int lf_batch_info(const lf_ctx* c, uint32_t* nScenes, size_t* offsets) {
  if (!c) return LF_ERR_ARG;
  const uint32_t S = c->batch ? uint32_t(c->sceneParams.size()) : 1u;
  if (nScenes) *nScenes = S;
  if (offsets) {
    if (c->batch)
      for (uint32_t s = 0; s <= S; ++s) offsets[s] = c->sceneBegin[s];
    else {
      offsets[0] = 0;
      offsets[1] = c->n;
    }
  }
  return LF_OK;
}

int lf_batch_select(lf_ctx* c, uint32_t scene) {
  if (!c) return LF_ERR_ARG;
  const uint32_t S = c->batch ? uint32_t(c->sceneParams.size()) : 1u;
  if (scene >= S) return fail(c, LF_ERR_ARG, "lf_batch_select: no such scene");
  c->mcScene = int(scene);
  return LF_OK;
}

int lf_exec_configure(lf_ctx* c, const lf_exec_params* e) {
  if (!c || !e) return LF_ERR_ARG;
  if (e->threads != 128 && e->threads != 256 && e->threads != 512)
    return fail(c, LF_ERR_ARG, "lf_exec_configure: threads must be 128, 256 or 512");
  if (!(e->skin >= 0.0f && e->skin <= 1.0f))
    return fail(c, LF_ERR_ARG, "lf_exec_configure: skin must be in [0, 1] (fraction of h)");
  if (e->particles_per_thread != 1 && e->particles_per_thread != 2 && e->particles_per_thread != 4)
    return fail(c, LF_ERR_ARG, "lf_exec_configure: particles_per_thread must be 1, 2 or 4");
  if (e->mc_tile >= lf::kMcTiles)
    return fail(c, LF_ERR_ARG, "lf_exec_configure: mc_tile must be 0, 1, 2 or 3");
  lf::dropGraph(c);
  c->chunkThreads = e->threads;
  c->persistent = e->persistent != 0;
  c->skinFrac = e->skin;
  c->chunkPpt = e->particles_per_thread;
  c->mcTile = e->mc_tile;
  c->gridValid = false;  // constants fold the skin: the next call must step or build the grid
  if (!c->batch && c->params.h > 0.0f) (void)lf::setParams(c, &c->params);
  return LF_OK;
}

int lf_exec_config(lf_ctx* c, uint32_t threads, int persistent, float skin) {
  if (!c) return LF_ERR_ARG;
  const lf_exec_params e{threads, persistent, skin, c->chunkPpt, c->mcTile};
  return lf_exec_configure(c, &e);
}

int lf_autotune(lf_ctx* c, uint32_t steps, lf_tune_result* out) {
  if (!c || steps == 0) return LF_ERR_ARG;
  c->posEventValid = false;  // P and the table are rewritten outside the event
  if (c->n == 0 || !(c->params.h > 0.0f))
    return fail(c, LF_ERR_STATE, "lf_autotune: needs a state and one step (for its parameters)");
  // trial steps would append and erase particles: the state could not be put back
  if (!c->sources.empty() || !c->hostDrains.empty())
    return fail(c, LF_ERR_STATE, "lf_autotune: sources and drains change the particle count");
  LF_HIP(c, hipSetDevice(c->device));
  const size_t n = c->n;
  const int cur0 = c->cur;
  // snapshot of the state
  float4 *x0 = nullptr, *v0 = nullptr;
  uint8_t* t0 = nullptr;
  uint32_t* id0 = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = LF_OK;
  auto cleanup = [&]() {
    if (x0) (void)hipFree(x0);
    if (v0) (void)hipFree(v0);
    if (t0) (void)hipFree(t0);
    if (id0) (void)hipFree(id0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  auto hipOk = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == LF_OK) rc = lf::hipFail(c, e, what);
    return rc == LF_OK;
  };
  const bool ok = hipOk(hipMalloc(reinterpret_cast<void**>(&x0), n * sizeof(float4)), "autotune") &&
                  hipOk(hipMalloc(reinterpret_cast<void**>(&v0), n * sizeof(float4)), "autotune") &&
                  hipOk(hipMalloc(reinterpret_cast<void**>(&t0), n), "autotune") &&
                  hipOk(hipMalloc(reinterpret_cast<void**>(&id0), n * sizeof(uint32_t)), "autotune") &&
                  hipOk(hipEventCreate(&e0), "autotune") && hipOk(hipEventCreate(&e1), "autotune");
  if (!ok) {
    cleanup();
    return rc;
  }
  auto copyState = [&](bool save) {
    const size_t b4 = n * sizeof(float4);
    if (save) {
      return hipOk(hipMemcpyAsync(x0, c->X[cur0].p, b4, hipMemcpyDeviceToDevice, c->stream), "autotune") &&
             hipOk(hipMemcpyAsync(v0, c->V[cur0].p, b4, hipMemcpyDeviceToDevice, c->stream), "autotune") &&
             hipOk(hipMemcpyAsync(t0, c->T[cur0].p, n, hipMemcpyDeviceToDevice, c->stream), "autotune") &&
             hipOk(hipMemcpyAsync(id0, c->ID[cur0].p, n * 4, hipMemcpyDeviceToDevice, c->stream), "autotune");
    }
    c->cur = cur0;
    return hipOk(hipMemcpyAsync(c->X[cur0].p, x0, b4, hipMemcpyDeviceToDevice, c->stream), "autotune") &&
           hipOk(hipMemcpyAsync(c->V[cur0].p, v0, b4, hipMemcpyDeviceToDevice, c->stream), "autotune") &&
           hipOk(hipMemcpyAsync(c->T[cur0].p, t0, n, hipMemcpyDeviceToDevice, c->stream), "autotune") &&
           hipOk(hipMemcpyAsync(c->ID[cur0].p, id0, n * 4, hipMemcpyDeviceToDevice, c->stream), "autotune");
  };
  if (!copyState(true)) {
    cleanup();
    return rc;
  }
  const bool prof = c->prof;
  c->prof = false;
  // trial steps leave the last step's query answers alone
  const uint32_t nQueries = c->nQueries;
  c->nQueries = 0;
  const uint32_t threadsList[3] = {128, 256, 512};
  const float skins[4] = {0.1f, 0.15f, 0.2f, 0.3f};
  // (particles per thread, persistent): the queue takes chunks one at a time
  const uint32_t pptPers[4][2] = {{1, 0}, {2, 0}, {4, 0}, {1, 1}};
  lf_tune_result best{};
  best.ms_per_step = INFINITY;
  best.mc_tile = c->mcTile;
  uint32_t trials = 0;
  const uint32_t tile0 = c->mcTile;
  for (uint32_t th : threadsList)
    for (int pp = 0; pp < 4 && rc == LF_OK; ++pp)
      for (float sk : skins) {
        const lf_exec_params ex{th, int32_t(pptPers[pp][1]), sk, pptPers[pp][0], tile0};
        if ((rc = lf_exec_configure(c, &ex))) break;
        if (!copyState(false)) break;
        if ((rc = lf::pbfStepOnce(c))) break;  // warm-up (allocations, caches)
        if (!hipOk(hipEventRecord(e0, c->stream), "autotune")) break;
        for (uint32_t s = 0; s < steps && rc == LF_OK; ++s) rc = lf::pbfStepOnce(c);
        if (rc) break;
        if (!hipOk(hipEventRecord(e1, c->stream), "autotune") ||
            !hipOk(hipEventSynchronize(e1), "autotune"))
          break;
        float ms = 0.0f;
        if (!hipOk(hipEventElapsedTime(&ms, e0, e1), "autotune")) break;
        ++trials;
        if (ms / float(steps) < best.ms_per_step) {
          best.threads = th;
          best.persistent = int32_t(pptPers[pp][1]);
          best.skin = sk;
          best.particles_per_thread = pptPers[pp][0];
          best.ms_per_step = ms / float(steps);
        }
      }
  if (rc == LF_OK) {
    const lf_exec_params ex{best.threads, best.persistent, best.skin, best.particles_per_thread,
                            tile0};
    rc = lf_exec_configure(c, &ex);
  }
  // the density field's tile shape, on the last density lattice sampled, from
  // one step of the restored state
  if (rc == LF_OK && c->haveDensity && copyState(false) && (rc = lf::pbfStepOnce(c)) == LF_OK) {
    const lf_mc_params m = c->lastDensity;
    float bestMc = INFINITY;
    for (uint32_t t = 0; t < lf::kMcTiles && rc == LF_OK; ++t) {
      c->mcTile = t;
      if ((rc = lf::mcSample(c, m))) break;  // warm-up
      if ((rc = lf::mcJoin(c))) break;
      if (!hipOk(hipEventRecord(e0, c->stream), "autotune")) break;
      for (int r = 0; r < 3 && rc == LF_OK; ++r) {
        rc = lf::mcSample(c, m);
        if (rc == LF_OK) rc = lf::mcJoin(c);
      }
      if (rc) break;
      if (!hipOk(hipEventRecord(e1, c->stream), "autotune") ||
          !hipOk(hipEventSynchronize(e1), "autotune"))
        break;
      float ms = 0.0f;
      if (!hipOk(hipEventElapsedTime(&ms, e0, e1), "autotune")) break;
      ++trials;
      if (ms / 3.0f < bestMc) {
        bestMc = ms / 3.0f;
        best.mc_tile = t;
        best.mc_ms = bestMc;
      }
    }
    c->mcTile = best.mc_tile;
  }
  c->prof = prof;
  c->nQueries = nQueries;
  if (rc == LF_OK) (void)copyState(false);
  if (rc == LF_OK) (void)hipOk(hipStreamSynchronize(c->stream), "autotune");
  cleanup();
  // the grid, lists and marching-cubes buffers hold trial data now
  c->gridValid = false;
  c->stepDone = false;
  c->mcSampled = c->mcCounted = c->mcEmitted = false;
  best.trials = trials;
  if (out) *out = best;
  return rc;
}

int lf_graph_capture(lf_ctx* c, uint32_t nsteps, const lf_mc_params* m) {
  if (!c) return LF_ERR_ARG;
  c->posEventValid = false;  // P and the table are rewritten outside the event
  lf::syncStreams(c);  // a capture must not wait on work recorded before it
  if (nsteps == 0 || (nsteps & 1u))
    return fail(c, LF_ERR_ARG, "lf_graph_capture: nsteps must be even and > 0");
  if (!c->stepDone || c->n == 0)
    return fail(c, LF_ERR_STATE, "lf_graph_capture: run one eager step with these sizes first");
  if (!c->sources.empty() || !c->hostDrains.empty())
    return fail(c, LF_ERR_STATE, "lf_graph_capture: sources and drains change the particle count");
  LF_HIP(c, hipSetDevice(c->device));
  lf::dropGraph(c);
  int rc = lf::prepareScenes(c);  // uploads now, not inside the capture
  if (rc) return rc;
  LF_HIP(c, hipStreamSynchronize(c->stream));
  const bool prof = c->prof;
  c->prof = false;
  c->capturing = true;
  hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    c->capturing = false;
    c->prof = prof;
    return lf::hipFail(c, e, "hipStreamBeginCapture");
  }
  for (uint32_t s = 0; s < nsteps && rc == LF_OK; ++s) {
    rc = lf::pbfStepOnce(c);
    if (rc == LF_OK && m) rc = lf::mcSample(c, *m);
    if (rc == LF_OK && m) rc = lf::mcClassifyCount(c, *m);
    if (rc == LF_OK && m) rc = lf::mcEmit(c, *m);
  }
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(c->stream, &g);
  c->capturing = false;
  c->prof = prof;
  if (rc == LF_OK && e != hipSuccess) rc = lf::hipFail(c, e, "hipStreamEndCapture");
  if (rc == LF_OK) {
    e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      c->graph = nullptr;
      rc = lf::hipFail(c, e, "hipGraphInstantiate");
    }
  }
  if (g) (void)hipGraphDestroy(g);
  if (rc == LF_OK) c->graphSteps = int(nsteps);
  return rc;
}

int lf_graph_launch(lf_ctx* c, uint32_t times) {
  if (!c) return LF_ERR_ARG;
  c->posEventValid = false;  // P and the table are rewritten outside the event
  if (lf::mcJoin(c)) return lf::fail(c, LF_ERR_HIP, "waiting for the marching-cubes stream");
  if (!c->graph) return fail(c, LF_ERR_STATE, "lf_graph_launch: no captured graph");
  LF_HIP(c, hipSetDevice(c->device));
  int rc = lf::checkHealth(c, false);
  if (rc) return rc;
  for (uint32_t t = 0; t < times; ++t) LF_HIP(c, hipGraphLaunch(c->graph, c->stream));
  return LF_OK;
}

int lf_diagnostics(lf_ctx* c, uint32_t scene, lf_diag* out) {
  if (!c || !out) return LF_ERR_ARG;
  if (!c->gridValid) return fail(c, LF_ERR_STATE, "lf_diagnostics: no step or grid yet");
  const uint32_t S = uint32_t(c->sceneConsts.size());
  if (scene >= S) return fail(c, LF_ERR_ARG, "lf_diagnostics: no such scene");
  LF_HIP(c, hipSetDevice(c->device));
  int rc = lf::diagState(c);
  if (rc) return rc;
  std::vector<double> st(lf::kStatSlots * lf::kSumWidth);
  uint32_t passes = 0;
  LF_HIP(c, hipMemcpyAsync(st.data(), c->sceneStats.p + size_t(scene) * st.size(),
                           st.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  LF_HIP(c, hipMemcpyAsync(&passes, c->scenePasses.p + scene, 4, hipMemcpyDeviceToHost, c->stream));
  LF_HIP(c, hipStreamSynchronize(c->stream));
  std::memset(out, 0, sizeof(*out));
  const double* s = st.data() + size_t(lf::kPassSlots) * lf::kSumWidth;
  out->fluid = uint64_t(s[0]);
  out->mass = s[1];
  out->momentum[0] = s[2];
  out->momentum[1] = s[3];
  out->momentum[2] = s[4];
  out->density_error_mean = s[0] > 0.0 ? s[5] / s[0] : 0.0;
  out->density_error_max = s[6];
  out->passes = c->sceneConsts[scene].densityTol > 0.0f ? passes : c->params.iterations;
  for (int i = 0; i < lf::kPassSlots; ++i) out->pass_error[i] = st[size_t(i) * lf::kSumWidth + 5];
  return LF_OK;
}

/* Keep marching-cubes case indices for lf_read(LF_ARR_CASES) (parity tests). */
int lf_mc_keep_cases(lf_ctx* c, int on) {
  if (!c) return LF_ERR_ARG;
  c->keepCases = on != 0;
  return LF_OK;
}

}  // extern "C"
