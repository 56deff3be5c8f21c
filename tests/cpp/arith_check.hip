// The pair terms' shortened correctly rounded square root and division
// (sqrtRn, divRn in libfluid_amd/csrc/lf_internal.hpp) against hipcc's own
// sqrtf and FP32 division, bitwise, on the GPU:
//  * sqrtRn(x) == sqrtf(x) for every float x >= 2^-96 (1.88e9 values);
//  * sqrtRn(x) <= 2^-47 for every float 0 <= x < 2^-96 (the values for which
//    the pair terms only need r < 1e-8 and r * r negligible next to h * h);
//  * divRn(a, b) == a / b for 2^32 spiky-term operands a = (h - r)^2, b = r
//    with 1e-4 <= h <= 1e4 and 1e-8 <= r <= h (uniform, log-uniform and
//    within 1024 ulp of h), and for 2^32 random operands with exponents in
//    [-60, 60].
// Prints one line per check and exits 0 when nothing differs.
#include <cstdio>

#include "lf_internal.hpp"

namespace {

__device__ inline uint64_t mix(uint64_t z) {  // splitmix64 finaliser
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ inline float u01(uint64_t& s) {
  s = mix(s);
  return float(s >> 40) * 0x1p-24f;
}
__device__ inline void miss(unsigned long long* bad, uint32_t* first, uint32_t a, uint32_t b) {
  if (atomicAdd(bad, 1ull) == 0ull) {
    first[0] = a;
    first[1] = b;
  }
}

__global__ void k_sqrt(uint32_t lo, uint32_t hi, unsigned long long* bad, uint32_t* first) {
  const uint64_t n = uint64_t(hi) - lo;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const float x = __uint_as_float(lo + uint32_t(i));
    if (__float_as_uint(lf::sqrtRn(x)) != __float_as_uint(__builtin_sqrtf(x)))
      miss(bad, first, __float_as_uint(x), 0u);
  }
}

__global__ void k_sqrt_small(uint32_t hi, unsigned long long* bad, uint32_t* first) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < hi;
       i += uint64_t(gridDim.x) * blockDim.x) {
    const float x = __uint_as_float(uint32_t(i));
    // unsigned compare: also rejects negative results and NaN
    if (__float_as_uint(lf::sqrtRn(x)) > __float_as_uint(0x1p-47f))
      miss(bad, first, __float_as_uint(x), 0u);
  }
}

// mode 0: spiky-term operands; mode 1: random operands, exponents in [-60, 60]
__global__ void k_div(int mode, uint64_t n, uint64_t seed, unsigned long long* bad,
                      uint32_t* first) {
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x) {
    uint64_t s = mix(seed ^ (i * 0x2545F4914F6CDD1Dull));
    float a, b;
    if (mode == 0) {
      const float h = __builtin_exp2f(-13.28f + 26.56f * u01(s));  // 1e-4 .. 1e4
      const float v = u01(s);
      const uint32_t kind = uint32_t(s >> 13) & 3u;
      float r;
      if (kind == 0)
        r = h * v;
      else if (kind == 1)
        r = h * __builtin_exp2f(-40.0f * v);
      else if (kind == 2)
        r = __uint_as_float(__float_as_uint(h) - (uint32_t(s >> 20) & 1023u));
      else
        r = h * (1.0f - v * 0x1p-10f);
      if (!(r >= 1e-8f) || !(h >= 1e-4f) || !(h <= 1e4f) || r > h) continue;
      const float q = h - r;
      a = q * q;
      b = r;
    } else {
      s = mix(s);
      const uint32_t ea = 127u - 60u + uint32_t(s % 121u);
      const uint32_t eb = 127u - 60u + uint32_t((s >> 8) % 121u);
      a = __uint_as_float((ea << 23) | (uint32_t(s >> 20) & 0x7fffffu));
      b = __uint_as_float((eb << 23) | (uint32_t(s >> 43) & 0x7fffffu));
    }
    if (__float_as_uint(lf::divRn(a, b)) != __float_as_uint(a / b))
      miss(bad, first, __float_as_uint(a), __float_as_uint(b));
  }
}

}  // namespace

int main() {
  unsigned long long* bad = nullptr;
  uint32_t* first = nullptr;
  if (hipMalloc(&bad, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&first, 8 * sizeof(uint32_t)) != hipSuccess) {
    std::printf("hipMalloc failed\n");
    return 2;
  }
  if (hipMemset(bad, 0, 4 * sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(first, 0, 8 * sizeof(uint32_t)) != hipSuccess) {
    std::printf("hipMemset failed\n");
    return 2;
  }
  const uint32_t lo = __builtin_bit_cast(uint32_t, 0x1p-96f), hi = 0x7f800000u;  // [2^-96, inf)
  const uint64_t n = 1ull << 32;
  hipLaunchKernelGGL(k_sqrt, dim3(4096), dim3(256), 0, 0, lo, hi, bad + 0, first + 0);
  hipLaunchKernelGGL(k_sqrt_small, dim3(4096), dim3(256), 0, 0, lo, bad + 1, first + 2);
  hipLaunchKernelGGL(k_div, dim3(4096), dim3(256), 0, 0, 0, n, 12345ull, bad + 2, first + 4);
  hipLaunchKernelGGL(k_div, dim3(4096), dim3(256), 0, 0, 1, n, 999ull, bad + 3, first + 6);
  if (hipDeviceSynchronize() != hipSuccess) {
    std::printf("kernel error\n");
    return 3;
  }
  unsigned long long h[4];
  uint32_t f[8];
  if (hipMemcpy(h, bad, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(f, first, sizeof(f), hipMemcpyDeviceToHost) != hipSuccess) {
    std::printf("hipMemcpy failed\n");
    return 3;
  }
  std::printf("sqrt x >= 2^-96: %u values, mismatches %llu (first x=%08x)\n", hi - lo, h[0], f[0]);
  std::printf("sqrt x < 2^-96: %u values, above 2^-47 %llu (first x=%08x)\n", lo, h[1], f[2]);
  std::printf("div spiky operands: %llu samples, mismatches %llu (first a=%08x b=%08x)\n",
              (unsigned long long)n, h[2], f[4], f[5]);
  std::printf("div exponents -60..60: %llu samples, mismatches %llu (first a=%08x b=%08x)\n",
              (unsigned long long)n, h[3], f[6], f[7]);
  return (h[0] | h[1] | h[2] | h[3]) ? 1 : 0;
}
