// Device helpers shared by the HIP translation units of liblfric_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace lfric {
namespace {

constexpr int kBlock = 256;

inline int grid_for(long long n, int block = kBlock, long long cap = 1 << 30) {
  long long g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<int>(g);
}

// ---------------------------------------------------------------------------------------
// deterministic reductions: wavefront shuffles, then one partial per block; a one-block
// finishing kernel adds the partials in block order
// ---------------------------------------------------------------------------------------
template <int NV>
__device__ inline void wave_sum(double (&v)[NV]) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] += __shfl_down(v[q], off, 64);
}

template <int NV>
__device__ inline void block_sum(double (&v)[NV], double (&out)[NV]) {
  __shared__ double sh[kBlock / 64][NV];
  wave_sum<NV>(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[wid][q] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) s += sh[w][q];
    out[q] = s;
  }
  __syncthreads();
}

// Every block of a reduction kernel writes its partial sums; a one-block finishing kernel
// (k_finish) adds them in block order, so sums are deterministic run to run.
template <int NV>
__device__ inline void block_partials(double (&v)[NV], double *partials) {
  double bs[NV];
  block_sum<NV>(v, bs);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) partials[blockIdx.x * NV + q] = bs[q];
}

template <typename T>
__device__ inline double to_d(T v) {
  return static_cast<double>(v);
}

__device__ inline size_t aidx(size_t gs, long long e, int p) {
  return static_cast<size_t>(e >> 6) * gs + static_cast<size_t>(p) * 64 + static_cast<size_t>(e & 63);
}

}  // namespace
}  // namespace lfric
