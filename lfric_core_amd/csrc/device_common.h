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
// (k_finish) adds them in block order, so sums are deterministic run to run.  Kernels whose
// blocks work on a remapped block id (xcd_block) pass that id, so the order of the sum follows
// the data and not the launch's XCD window.
template <int NV>
__device__ inline void block_partials(double (&v)[NV], double *partials, int block) {
  double bs[NV];
  block_sum<NV>(v, bs);
  if (threadIdx.x == 0)
#pragma unroll
    for (int q = 0; q < NV; ++q) partials[static_cast<size_t>(block) * NV + q] = bs[q];
}
template <int NV>
__device__ inline void block_partials(double (&v)[NV], double *partials) {
  block_partials<NV>(v, partials, static_cast<int>(blockIdx.x));
}

template <typename T>
__device__ inline double to_d(T v) {
  return static_cast<double>(v);
}

// Workgroups are handed to the 8 XCDs (each with its own L2) in turn by workgroup id.  This
// maps the workgroup id to a logical block so that, within each window of 8K blocks, each XCD
// gets K consecutive blocks: neighbouring blocks, which read the same x entries, then share an
// L2, while all XCDs still stream nearby addresses.  K <= 0 keeps dispatch order; the last
// partial window keeps dispatch order.  A bijection on [0, G).
__device__ inline int xcd_block(int b, int G, int K) {
  constexpr int kXcd = 8;
  if (K <= 0) return b;
  const int W = kXcd * K;
  const int base = (b / W) * W;
  if (base + W > G) return b;
  const int r = b - base;
  return base + (r % kXcd) * K + r / kXcd;
}

// Entries of an operator element per 16-byte chunk (4 FP32, 2 FP64).
template <typename T>
constexpr int chunk_entries() {
  return 16 / static_cast<int>(sizeof(T));
}
// Group stride of an operator whose elements hold n1*n2 entries: 64 elements per group,
// entries padded to whole chunks.
template <typename T>
constexpr size_t group_stride(int n1, int n2) {
  return static_cast<size_t>(64) * ((n1 * n2 + chunk_entries<T>() - 1) / chunk_entries<T>()) * chunk_entries<T>();
}
// Address of entry p of element e: group e/64, then chunk p/V, then element e%64, then p%V,
// so that one lane's V consecutive entries form one 16-byte load and a wavefront's 64 loads of
// one chunk form one contiguous 1 KiB block.
template <typename T>
__device__ inline size_t aidx(size_t gs, long long e, int p) {
  constexpr int V = chunk_entries<T>();
  return static_cast<size_t>(e >> 6) * gs + (static_cast<size_t>(p / V) * 64 + static_cast<size_t>(e & 63)) * V +
         static_cast<size_t>(p % V);
}


// Visit the entries of row r of element e in df2 order, f(j, a_rj), reading the row as whole
// 16-byte chunks.  A row that does not start on a chunk boundary (ndf2 not a multiple of the
// chunk) starts S entries into its first chunk; S is made a compile-time constant by a switch
// so that every index stays static.
template <typename T, int N2, int S, class F>
__device__ inline void for_row_s(const T *__restrict__ Ae, int c0, F &&f) {
  constexpr int V = chunk_entries<T>();
  constexpr int NCH = (N2 + S + V - 1) / V;
  typedef T VecT __attribute__((ext_vector_type(V)));
#pragma unroll
  for (int c = 0; c < NCH; ++c) {
    const VecT v = *reinterpret_cast<const VecT *>(Ae + static_cast<size_t>(c0 + c) * 64 * V);
#pragma unroll
    for (int q = 0; q < V; ++q) {
      const int j = c * V + q - S;
      if (j >= 0 && j < N2) f(j, v[q]);
    }
  }
}

template <typename T, int N2, class F>
__device__ inline void for_row(const T *__restrict__ A, size_t gs, long long e, int r, F &&f) {
  constexpr int V = chunk_entries<T>();
  const T *Ae = A + static_cast<size_t>(e >> 6) * gs + static_cast<size_t>(e & 63) * V;
  const int p0 = r * N2;
  if constexpr (N2 % V == 0) {
    for_row_s<T, N2, 0>(Ae, p0 / V, f);
  } else if constexpr (V == 2) {
    if (p0 & 1)
      for_row_s<T, N2, 1>(Ae, p0 / V, f);
    else
      for_row_s<T, N2, 0>(Ae, p0 / V, f);
  } else {
    switch (p0 & 3) {
      case 0: for_row_s<T, N2, 0>(Ae, p0 / V, f); break;
      case 1: for_row_s<T, N2, 1>(Ae, p0 / V, f); break;
      case 2: for_row_s<T, N2, 2>(Ae, p0 / V, f); break;
      default: for_row_s<T, N2, 3>(Ae, p0 / V, f); break;
    }
  }
}

}  // namespace
}  // namespace lfric
