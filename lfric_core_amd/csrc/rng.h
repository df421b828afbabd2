// Counter-based generator of the synthetic benchmark inputs, evaluated on the GPU.
//
//   bits(seed, stream, ctr) = mix(mix(seed + G*(stream+1)) ^ (ctr*K1 + K2)),
//     mix = SplitMix64 finaliser (Steele, Lea & Flood, OOPSLA 2014)
//   uniform(0,1) = ((bits >> 11) + 0.5) * 2^-53
//   normal(0,1)  = Marsaglia polar method on attempt pairs (ctr*64 + 2a, ctr*64 + 2a + 1)
//   log / exp    = only +,-,*,/ and fma (atanh series on [sqrt(1/2), sqrt(2)); Cody-Waite
//                  reduction and a degree-13 Taylor polynomial), so host and device round
//                  identically and every element matrix is the same on CPU and GPU.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace lfric {
namespace rng {

#pragma clang fp contract(off)

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t bits(uint64_t seed, uint64_t stream, uint64_t ctr) {
  const uint64_t k = mix64(seed + 0x9E3779B97F4A7C15ULL * (stream + 1));
  return mix64(k ^ (ctr * 0xD1B54A32D192ED03ULL + 0x8CB92BA72F3D8DD7ULL));
}

__host__ __device__ inline double uniform(uint64_t seed, uint64_t stream, uint64_t ctr) {
  return (static_cast<double>(bits(seed, stream, ctr) >> 11) + 0.5) * 0x1p-53;
}

__host__ __device__ inline double from_bits(uint64_t b) {
  union {
    uint64_t u;
    double d;
  } v;
  v.u = b;
  return v.d;
}
__host__ __device__ inline uint64_t to_bits(double d) {
  union {
    uint64_t u;
    double d;
  } v;
  v.d = d;
  return v.u;
}

constexpr double kLn2Hi = 0x1.62e42fefa3800p-1;
constexpr double kLn2Lo = 0x1.ef35793c76730p-45;

__host__ __device__ inline double log_pos(double x) {
  const uint64_t b = to_bits(x);
  int e = static_cast<int>((b >> 52) & 0x7ff) - 1023;
  double m = from_bits((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
  if (m > 0x1.6a09e667f3bcdp+0) {
    m = m * 0.5;
    e += 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  // 1/(2j+1), j = 12 .. 0
  double p = 1.0 / 25.0;
  p = fma(p, z, 1.0 / 23);
  p = fma(p, z, 1.0 / 21);
  p = fma(p, z, 1.0 / 19);
  p = fma(p, z, 1.0 / 17);
  p = fma(p, z, 1.0 / 15);
  p = fma(p, z, 1.0 / 13);
  p = fma(p, z, 1.0 / 11);
  p = fma(p, z, 1.0 / 9);
  p = fma(p, z, 1.0 / 7);
  p = fma(p, z, 1.0 / 5);
  p = fma(p, z, 1.0 / 3);
  p = fma(p, z, 1.0);
  const double lm = (2.0 * s) * p;
  const double de = static_cast<double>(e);
  return fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
}

__host__ __device__ inline double exp_small(double x) {
  const double n = rint(x * 0x1.71547652b82fep+0);
  double r = fma(-n, kLn2Hi, x);
  r = fma(-n, kLn2Lo, r);
  double p = 1.0 / 6227020800.0;
  p = fma(p, r, 1.0 / 479001600.0);
  p = fma(p, r, 1.0 / 39916800.0);
  p = fma(p, r, 1.0 / 3628800.0);
  p = fma(p, r, 1.0 / 362880.0);
  p = fma(p, r, 1.0 / 40320.0);
  p = fma(p, r, 1.0 / 5040.0);
  p = fma(p, r, 1.0 / 720.0);
  p = fma(p, r, 1.0 / 120.0);
  p = fma(p, r, 1.0 / 24.0);
  p = fma(p, r, 1.0 / 6.0);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  const int ni = static_cast<int>(n);
  return p * from_bits(static_cast<uint64_t>(ni + 1023) << 52);
}

__host__ __device__ inline double normal(uint64_t seed, uint64_t stream, uint64_t ctr) {
  for (uint64_t a = 0; a < 32; ++a) {
    const double v1 = 2.0 * uniform(seed, stream, ctr * 64 + 2 * a) - 1.0;
    const double v2 = 2.0 * uniform(seed, stream, ctr * 64 + 2 * a + 1) - 1.0;
    const double s = fma(v1, v1, v2 * v2);
    if (s < 1.0 && s > 0.0) return v1 * sqrt(-2.0 * log_pos(s) / s);
  }
  return 0.0;
}

// streams of the synthetic inputs
constexpr uint64_t kStreamB = 1;      // B_e entries
constexpr uint64_t kStreamScale = 2;  // per-column exponents of the scaled operator
constexpr uint64_t kStreamRhs = 3;    // right-hand sides

#pragma clang fp contract(on)

}  // namespace rng
}  // namespace lfric
