// PBF constraint passes on gfx950: density/lambda, position correction
// (Jacobi), finalise, XSPH viscosity and colour diffusion.
//
// One work-item per Morton-sorted particle; candidates come from the 27 cells
// around the particle's own cell in the reference's order
// (oclsph_kernel.h:150-192), so each sum runs in the same order as the
// reference and the CPU restatement.
//
// Buffers (sorted order): P = (p*, mass), PL = (p*, lambda), X = (world
// position, mass), V = (velocity, colour bits), T = type.
#include "lf_internal.hpp"

namespace lf {
namespace {

constexpr int kThreads = 256;

// sph_lambda (oclsph_kernel.h:305-332). KAHAN: Neumaier-compensated density
// and gradient sums (mixed-precision option).
template <bool KAHAN>
__global__ __launch_bounds__(kThreads) void k_lambda(uint32_t n, PbfConsts c,
                                                     const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ table,
                                                     const uint8_t* __restrict__ T,
                                                     const float4* __restrict__ P,
                                                     float4* __restrict__ PL,
                                                     float* __restrict__ rhoOut) {
  const uint32_t a = blockIdx.x * kThreads + threadIdx.x;
  if (a >= n) return;
  const float4 pa = P[a];
  if (T[a] == LF_OBSTACLE) {
    PL[a] = make_float4(pa.x, pa.y, pa.z, 0.0f);
    rhoOut[a] = 0.0f;
    return;
  }
  float gx = 0.0f, gy = 0.0f, gz = 0.0f, rho = 0.0f;
  float cgx = 0.0f, cgy = 0.0f, cgz = 0.0f, crho = 0.0f;
  forEachNeighbourCell(keys[a], table, c.G, [&](uint32_t s, uint32_t e) {
    for (uint32_t b = s; b < e; ++b) {
      const float4 pb = P[b];
      const float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
      const float d2 = dist2(dx, dy, dz);
      if (d2 <= c.rmax2) {
        const float r = __builtin_sqrtf(d2);
        const float w = poly6W(c, r);
        const float sc = spikyScale(c, r);
        if (!KAHAN) {
          gx = __builtin_fmaf(dx * sc, c.rhoRecip, gx);
          gy = __builtin_fmaf(dy * sc, c.rhoRecip, gy);
          gz = __builtin_fmaf(dz * sc, c.rhoRecip, gz);
          rho = __builtin_fmaf(pa.w, w, rho);
        } else {
          const float tg[3] = {(dx * sc) * c.rhoRecip, (dy * sc) * c.rhoRecip,
                               (dz * sc) * c.rhoRecip};
          float* acc[3] = {&gx, &gy, &gz};
          float* cmp[3] = {&cgx, &cgy, &cgz};
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            const float sum = *acc[k] + tg[k];
            *cmp[k] += (fabsf(*acc[k]) >= fabsf(tg[k])) ? (*acc[k] - sum) + tg[k]
                                                       : (tg[k] - sum) + *acc[k];
            *acc[k] = sum;
          }
          const float term = pa.w * w;
          const float sum = rho + term;
          crho += (fabsf(rho) >= fabsf(term)) ? (rho - sum) + term : (term - sum) + rho;
          rho = sum;
        }
      }
    }
  });
  if (KAHAN) {
    gx += cgx;
    gy += cgy;
    gz += cgz;
    rho += crho;
  }
  const float norm2 = dist2(gx, gy, gz);
  const float C1 = rho / c.rho0 - 1.0f;
  PL[a] = make_float4(pa.x, pa.y, pa.z, -C1 / (norm2 + c.cfm));
  rhoOut[a] = rho;
}

// sph_delta (oclsph_kernel.h:335-384) without colliders; Jacobi: reads PL,
// writes Pout. deltaP is not stored (nothing reads it after the pass).
__global__ __launch_bounds__(kThreads) void k_delta(uint32_t n, PbfConsts c,
                                                    const uint32_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ table,
                                                    const uint8_t* __restrict__ T,
                                                    const float4* __restrict__ PL,
                                                    const float4* __restrict__ X,
                                                    float4* __restrict__ Pout) {
  const uint32_t a = blockIdx.x * kThreads + threadIdx.x;
  if (a >= n) return;
  const float4 pa = PL[a];
  const float mass = X[a].w;
  if (T[a] == LF_OBSTACLE) {
    Pout[a] = make_float4(pa.x, pa.y, pa.z, mass);
    return;
  }
  const float la = pa.w;
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  forEachNeighbourCell(keys[a], table, c.G, [&](uint32_t s, uint32_t e) {
    for (uint32_t b = s; b < e; ++b) {
      const float4 pb = PL[b];
      const float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
      const float d2 = dist2(dx, dy, dz);
      if (d2 <= c.rmax2) {
        const float r = __builtin_sqrtf(d2);
        const float q = poly6W(c, r) / c.wdq;
        const float corr = -c.corrK * ((q * q) * (q * q));
        const float factor = ((la + pb.w) + corr) / c.rho0;
        const float sc = spikyScale(c, r);
        ax = __builtin_fmaf(dx * sc, factor, ax);
        ay = __builtin_fmaf(dy * sc, factor, ay);
        az = __builtin_fmaf(dz * sc, factor, az);
      }
    }
  });
  // pos = (p* + dp) * scale, clamped to the bounds (:364-371), p* = pos/scale
  const float acc[3] = {ax, ay, az};
  const float p[3] = {pa.x, pa.y, pa.z};
  float o[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    float pos = (p[k] + acc[k]) * c.scale;
    pos = (c.minB[k] < pos) ? pos : c.minB[k];
    pos = (pos < c.maxB[k]) ? pos : c.maxB[k];
    o[k] = pos / c.scale;
  }
  Pout[a] = make_float4(o[0], o[1], o[2], mass);
}

// sph_finalise (oclsph_kernel.h:404-416). Also sets P.w to the fluid mass (0
// for obstacles) for the density field, and moves the diffused colour into V.w.
__global__ __launch_bounds__(kThreads) void k_finalise(uint32_t n, PbfConsts c,
                                                       const uint8_t* __restrict__ T,
                                                       float4* __restrict__ P,
                                                       float4* __restrict__ X,
                                                       float4* __restrict__ V,
                                                       const uint32_t* __restrict__ diffused,
                                                       int useDiffused) {
  const uint32_t a = blockIdx.x * kThreads + threadIdx.x;
  if (a >= n) return;
  float4 p = P[a];
  if (T[a] == LF_OBSTACLE) {
    p.w = 0.0f;
    P[a] = p;
    return;
  }
  const float4 x = X[a];
  const float4 v = V[a];
  const float dX = p.x - x.x / c.scale;
  const float dY = p.y - x.y / c.scale;
  const float dZ = p.z - x.z / c.scale;
  X[a] = make_float4(p.x * c.scale, p.y * c.scale, p.z * c.scale, x.w);
  const float col = useDiffused ? __uint_as_float(diffused[a]) : v.w;
  V[a] = make_float4(__builtin_fmaf(dX, c.invDt, v.x) * c.vd,
                     __builtin_fmaf(dY, c.invDt, v.y) * c.vd,
                     __builtin_fmaf(dZ, c.invDt, v.z) * c.vd, col);
}

// XSPH viscosity (absent in the reference): v_a += (c/rho0) sum_b m_b (v_b -
// v_a) W6(r) over fluid b within h, Jacobi (Vin -> Vout). P.w is the fluid
// mass after finalise.
__global__ __launch_bounds__(kThreads) void k_xsph(uint32_t n, PbfConsts c,
                                                   const uint32_t* __restrict__ keys,
                                                   const uint32_t* __restrict__ table,
                                                   const uint8_t* __restrict__ T,
                                                   const float4* __restrict__ P,
                                                   const float4* __restrict__ Vin,
                                                   float4* __restrict__ Vout) {
  const uint32_t a = blockIdx.x * kThreads + threadIdx.x;
  if (a >= n) return;
  const float4 va = Vin[a];
  if (T[a] == LF_OBSTACLE) {
    Vout[a] = va;
    return;
  }
  const float4 pa = P[a];
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  forEachNeighbourCell(keys[a], table, c.G, [&](uint32_t s, uint32_t e) {
    for (uint32_t b = s; b < e; ++b) {
      const float4 pb = P[b];
      const float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
      const float d2 = dist2(dx, dy, dz);
      if (d2 <= c.rmax2 && T[b] != LF_OBSTACLE) {
        const float mw = pb.w * poly6W(c, __builtin_sqrtf(d2));
        const float4 vb = Vin[b];
        ax = __builtin_fmaf(mw, vb.x - va.x, ax);
        ay = __builtin_fmaf(mw, vb.y - va.y, ay);
        az = __builtin_fmaf(mw, vb.z - va.z, az);
      }
    }
  });
  Vout[a] = make_float4(__builtin_fmaf(c.xsphRho, ax, va.x), __builtin_fmaf(c.xsphRho, ay, va.y),
                        __builtin_fmaf(c.xsphRho, az, va.z), va.w);
}

__device__ inline void unpackARGB(uint32_t c, float out[4]) {
  out[0] = float((c >> 24) & 0xFFu);
  out[1] = float((c >> 16) & 0xFFu);
  out[2] = float((c >> 8) & 0xFFu);
  out[3] = float(c & 0xFFu);
}
// convert_uchar of a value already clamped to [8, 255]: round toward zero
__device__ inline uint32_t toU8(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 255.0f) return 255u;
  return uint32_t(v);
}

// sph_diffuse (oclsph_kernel.h:281-303): mean colour of the non-obstacle
// candidates, mix(c_a, mean * 1.33, dt/750), clamp [8, 255], truncate.
__global__ __launch_bounds__(kThreads) void k_diffuse(uint32_t n, PbfConsts c,
                                                      const uint32_t* __restrict__ keys,
                                                      const uint32_t* __restrict__ table,
                                                      const uint8_t* __restrict__ T,
                                                      const float4* __restrict__ V,
                                                      uint32_t* __restrict__ diffused) {
  const uint32_t a = blockIdx.x * kThreads + threadIdx.x;
  if (a >= n) return;
  const uint32_t ca = __float_as_uint(V[a].w);
  if (T[a] == LF_OBSTACLE) {
    diffused[a] = ca;
    return;
  }
  float mix[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  int N = 0;
  forEachNeighbourCell(keys[a], table, c.G, [&](uint32_t s, uint32_t e) {
    for (uint32_t b = s; b < e; ++b) {
      if (T[b] == LF_OBSTACLE) continue;
      float cb[4];
      unpackARGB(__float_as_uint(V[b].w), cb);
#pragma unroll
      for (int k = 0; k < 4; ++k) mix[k] += cb[k];
      ++N;
    }
  });
  float cf[4];
  unpackARGB(ca, cf);
  const float t = c.dt / 750.0f;
  uint32_t out = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const float y = (mix[k] / float(N)) * 1.33f;
    float o = __builtin_fmaf(y - cf[k], t, cf[k]);
    o = fminf(fmaxf(o, 8.0f), 255.0f);
    out = (out << 8) | toU8(o);
  }
  diffused[a] = out;
}

}  // namespace

int pbfLambda(Ctx* c) {
  const unsigned nb = blocksFor(c->n, kThreads);
  if (c->n == 0) return LF_OK;
  const uint32_t* keys = c->sortKey[c->keysIdx].p;
  if (c->consts.flags & LF_FLAG_KAHAN)
    return launch(c, "pbf_lambda_kahan", k_lambda<true>, dim3(nb), dim3(kThreads), 0,
                  uint32_t(c->n), c->consts, keys, (const uint32_t*)c->table.p,
                  (const uint8_t*)c->T[c->cur].p, (const float4*)c->P.p, c->PL.p, c->rho.p);
  return launch(c, "pbf_lambda", k_lambda<false>, dim3(nb), dim3(kThreads), 0, uint32_t(c->n),
                c->consts, keys, (const uint32_t*)c->table.p, (const uint8_t*)c->T[c->cur].p,
                (const float4*)c->P.p, c->PL.p, c->rho.p);
}

int pbfDelta(Ctx* c) {
  if (c->n == 0) return LF_OK;
  const unsigned nb = blocksFor(c->n, kThreads);
  return launch(c, "pbf_delta", k_delta, dim3(nb), dim3(kThreads), 0, uint32_t(c->n), c->consts,
                (const uint32_t*)c->sortKey[c->keysIdx].p, (const uint32_t*)c->table.p,
                (const uint8_t*)c->T[c->cur].p, (const float4*)c->PL.p,
                (const float4*)c->X[c->cur].p, c->P.p);
}

int pbfFinalise(Ctx* c) {
  if (c->n == 0) return LF_OK;
  const unsigned nb = blocksFor(c->n, kThreads);
  return launch(c, "pbf_finalise", k_finalise, dim3(nb), dim3(kThreads), 0, uint32_t(c->n),
                c->consts, (const uint8_t*)c->T[c->cur].p, c->P.p, c->X[c->cur].p,
                c->V[c->cur].p, (const uint32_t*)c->diffused.p,
                int((c->consts.flags & LF_FLAG_DIFFUSE) != 0));
}

int pbfXsph(Ctx* c) {
  if (c->n == 0 || c->consts.xsph == 0.0f) return LF_OK;
  const unsigned nb = blocksFor(c->n, kThreads);
  const int nxt = 1 - c->cur;
  int rc = ensure(c, c->V[nxt], c->n, "velocities");
  if (rc) return rc;
  rc = launch(c, "pbf_xsph", k_xsph, dim3(nb), dim3(kThreads), 0, uint32_t(c->n), c->consts,
              (const uint32_t*)c->sortKey[c->keysIdx].p, (const uint32_t*)c->table.p,
              (const uint8_t*)c->T[c->cur].p, (const float4*)c->P.p,
              (const float4*)c->V[c->cur].p, c->V[nxt].p);
  if (rc) return rc;
  std::swap(c->V[0], c->V[1]);  // the XSPH output becomes the current V
  return LF_OK;
}

int pbfDiffuse(Ctx* c) {
  if (c->n == 0) return LF_OK;
  const unsigned nb = blocksFor(c->n, kThreads);
  return launch(c, "pbf_diffuse", k_diffuse, dim3(nb), dim3(kThreads), 0, uint32_t(c->n),
                c->consts, (const uint32_t*)c->sortKey[c->keysIdx].p,
                (const uint32_t*)c->table.p, (const uint8_t*)c->T[c->cur].p,
                (const float4*)c->V[c->cur].p, c->diffused.p);
}

}  // namespace lf
