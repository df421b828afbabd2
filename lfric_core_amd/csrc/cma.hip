// Columnwise (CMA) operators and the vertical tridiagonal solve
// (components/science/source/kernel/columnwise_operators/*.F90,
// infrastructure/source/operator/columnwise_operator_mod.f90,
// components/science/source/kernel/solver/sci_tri_solve_kernel_mod.f90).
//
// A CMA operator maps between horizontally discontinuous spaces (W3, Wtheta, W2V-like: every
// dof of a cell column belongs to that column alone), so it is one banded matrix per column and
// every kernel here runs one thread per column with no colouring.  The bands are stored with
// the column index fastest, entry (row i, band b) of column pos at (i*bw + b)*ncols + pos, so a
// wavefront's 64 loads of one (row, band) are one contiguous 256-byte (FP32) segment.  The
// per-entry arithmetic is the reference's, in its order, compiled without contraction: each
// entry equals the reference kernel's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "device_common.h"
#include "kernels.h"

namespace lfric {

namespace {

struct CmaP {
  int alpha, beta, gm, gp, bw, nrow, ncol;
};

CmaP params_of(const lfric_cma &A) { return {A.alpha, A.beta, A.gamma_m, A.gamma_p, A.bw, A.nrow, A.ncol}; }

// Band limits of row i (1-based): j_minus = ceiling((alpha i - gamma_p) / beta), j_plus =
// floor((alpha i + gamma_m) / beta).  The reference divides in r_solver reals; for the small
// integers of lowest-order operators that equals exact integer division (beta > 0).
__host__ __device__ inline int floor_div(int a, int b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }
__host__ __device__ inline int jminus(const CmaP &p, int i) {
  return p.beta == 1 ? p.alpha * i - p.gp : -floor_div(-(p.alpha * i - p.gp), p.beta);
}
__host__ __device__ inline int jplus(const CmaP &p, int i) {
  return p.beta == 1 ? p.alpha * i + p.gm : floor_div(p.alpha * i + p.gm, p.beta);
}

__device__ inline size_t cidx(const CmaP &p, int nc, int i1, int b1, int pos) {
  return (static_cast<size_t>(i1 - 1) * p.bw + (b1 - 1)) * nc + pos;
}

#pragma clang fp contract(off)

// columnwise_op_asm_kernel_code (sci_columnwise_op_asm_kernel_mod.F90:96-140): zero the column,
// then add the element entry (df1, df2) of every layer at row banded_to(df1,k), band
// banded_from(df2,k) - j_minus + 1.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cma_asm(T *__restrict__ C, CmaP p, int nc, const T *__restrict__ A,
                                                    size_t gs, const int *__restrict__ ebase, uint32_t row_top,
                                                    int nl, int ndf_to, int ndf_from, const int *__restrict__ band_to,
                                                    const int *__restrict__ band_from) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos >= nc) return;
  for (int i = 1; i <= p.nrow; ++i)
    for (int b = 1; b <= p.bw; ++b) C[cidx(p, nc, i, b, pos)] = T(0);
  for (int k = 0; k < nl; ++k) {
    const long long el = ebase[pos] + k;
    for (int df1 = 0; df1 < ndf_to; ++df1) {
      const int i = band_to[k * ndf_to + df1];
      const int jm = jminus(p, i);
      const long long er = el + ((row_top >> df1) & 1u);
      for (int df2 = 0; df2 < ndf_from; ++df2) {
        const int j = band_from[k * ndf_from + df2];
        const size_t c = cidx(p, nc, i, j - jm + 1, pos);
        C[c] = C[c] + A[aidx<T>(gs, er, df1 * ndf_from + df2)];
      }
    }
  }
}

// Rows are walked in chunks of kRowChunk: the loads of a whole chunk (bands, right-hand side,
// indirection) are issued before its arithmetic, so a column's thread waits for memory once per
// chunk rather than once per row.  With one thread per column (55,296 at C96) the chip holds
// less than one wavefront per SIMD, and this load-level parallelism is what hides HBM latency.
constexpr int kRowChunk = 8;

// columnwise_op_app_kernel_code (sci_columnwise_op_app_kernel_mod.F90:107-131): row i of the
// column adds A(j - j_minus + 1, i) x(mu_j) for j from max(1, j_minus) to min(ncol, j_plus)
// onto lhs(mu_i), mu = column base + indirection.  With accumulate == 0 the sum starts from 0
// (setval_c(lhs, 0) before the kernel), else from lhs.  BW is the bandwidth when it is a
// compile-time 1, 2 or 3 (0: any).
template <typename T, int BW>
__global__ __launch_bounds__(kBlock) void k_cma_apply(const T *__restrict__ C, CmaP p, int nc,
                                                      const T *__restrict__ x, T *__restrict__ y,
                                                      const int *__restrict__ map_to, int ndf_to,
                                                      const int *__restrict__ map_from, int ndf_from,
                                                      const int *__restrict__ ind_to,
                                                      const int *__restrict__ ind_from, int accumulate) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos >= nc) return;
  const int bw = BW > 0 ? BW : p.bw;
  const size_t bt = static_cast<size_t>(map_to[static_cast<size_t>(pos) * ndf_to]);
  const size_t bf = static_cast<size_t>(map_from[static_cast<size_t>(pos) * ndf_from]);
  if constexpr (BW > 0) {
    for (int i0 = 1; i0 <= p.nrow; i0 += kRowChunk) {
      T a[kRowChunk][BW], xv[kRowChunk][BW], acc[kRowChunk];
      int jm[kRowChunk], j0[kRowChunk], j1[kRowChunk];
      size_t mu[kRowChunk];
#pragma unroll
      for (int u = 0; u < kRowChunk; ++u) {
        const int i = i0 + u;
        if (i > p.nrow) continue;
        mu[u] = bt + ind_to[i - 1];
        jm[u] = jminus(p, i);
        j0[u] = max(1, jm[u]);
        j1[u] = min(p.ncol, jplus(p, i));
        acc[u] = accumulate ? y[mu[u]] : T(0);
#pragma unroll
        for (int b = 0; b < BW; ++b) {
          const int j = jm[u] + b;
          const bool in = j >= j0[u] && j <= j1[u];
          a[u][b] = in ? C[cidx(p, nc, i, b + 1, pos)] : T(0);
          xv[u][b] = in ? x[bf + ind_from[j - 1]] : T(0);
        }
      }
#pragma unroll
      for (int u = 0; u < kRowChunk; ++u) {
        const int i = i0 + u;
        if (i > p.nrow) continue;
#pragma unroll
        for (int b = 0; b < BW; ++b) {
          const int j = jm[u] + b;
          if (j >= j0[u] && j <= j1[u]) acc[u] = acc[u] + a[u][b] * xv[u][b];
        }
        y[mu[u]] = acc[u];
      }
    }
  } else {
    for (int i = 1; i <= p.nrow; ++i) {
      const size_t mu_i = bt + ind_to[i - 1];
      const int jm = jminus(p, i), jp = jplus(p, i);
      T acc = accumulate ? y[mu_i] : T(0);
      for (int j = max(1, jm); j <= min(p.ncol, jp); ++j) acc = acc + C[cidx(p, nc, i, j - jm + 1, pos)] * x[bf + ind_from[j - 1]];
      y[mu_i] = acc;
    }
  }
  (void)bw;
}

// columnwise_op_appinv_kernel_code (sci_columnwise_op_appinv_kernel_mod.F90:107-146): Thomas
// algorithm on the tridiagonal column, bands (1, 2, 3) = (sub, diagonal, super).  The forward
// sweep keeps c' in the [nrow][ncols] work array and d' in lhs; the backward sweep finishes lhs.
// Both sweeps load a chunk of rows ahead of its arithmetic.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cma_appinv(const T *__restrict__ C, CmaP p, int nc,
                                                       const T *__restrict__ x, T *__restrict__ y,
                                                       const int *__restrict__ map, int ndf,
                                                       const int *__restrict__ ind, T *__restrict__ cw) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos >= nc) return;
  const size_t base = static_cast<size_t>(map[static_cast<size_t>(pos) * ndf]);
  const int nrow = p.nrow;
  T cprev = T(0), dprev = T(0);
  for (int i0 = 1; i0 <= nrow; i0 += kRowChunk) {
    T a1[kRowChunk], a2[kRowChunk], a3[kRowChunk], xv[kRowChunk];
    size_t mu[kRowChunk];
#pragma unroll
    for (int u = 0; u < kRowChunk; ++u) {
      const int i = i0 + u;
      if (i > nrow) continue;
      mu[u] = base + ind[i - 1];
      a1[u] = C[cidx(p, nc, i, 1, pos)];
      a2[u] = C[cidx(p, nc, i, 2, pos)];
      a3[u] = C[cidx(p, nc, i, 3, pos)];
      xv[u] = x[mu[u]];
    }
#pragma unroll
    for (int u = 0; u < kRowChunk; ++u) {
      const int i = i0 + u;
      if (i > nrow) continue;
      T inv, cp, dp;
      if (i == 1) {
        inv = T(1) / a2[u];
        cp = inv * a3[u];
        dp = inv * xv[u];
      } else {
        inv = T(1) / (a2[u] - a1[u] * cprev);
        cp = i < nrow ? inv * a3[u] : T(0);
        dp = inv * (xv[u] - a1[u] * dprev);
      }
      cw[static_cast<size_t>(i - 1) * nc + pos] = cp;
      y[mu[u]] = dp;
      cprev = cp;
      dprev = dp;
    }
  }
  T dnext = dprev;  // d'(nrow) is final
  for (int i0 = nrow - 1; i0 >= 1; i0 -= kRowChunk) {
    T cv[kRowChunk], dv[kRowChunk];
    size_t mu[kRowChunk];
#pragma unroll
    for (int u = 0; u < kRowChunk; ++u) {
      const int i = i0 - u;
      if (i < 1) continue;
      mu[u] = base + ind[i - 1];
      cv[u] = cw[static_cast<size_t>(i - 1) * nc + pos];
      dv[u] = y[mu[u]];
    }
#pragma unroll
    for (int u = 0; u < kRowChunk; ++u) {
      const int i = i0 - u;
      if (i < 1) continue;
      const T d = dv[u] - cv[u] * dnext;
      y[mu[u]] = d;
      dnext = d;
    }
  }
}

// LDS-staged forms for columns of up to kMaxStagedRows rows.  A block is one wavefront of 64
// columns.  The fields are column-contiguous (a column's dofs are consecutive), so reading them
// one lane per column touches 64 cache lines per instruction; instead the block copies its 64
// columns of x into LDS with lanes running along each column (64 consecutive values per
// instruction), the lanes then work one column each out of LDS, and the result goes back to
// HBM the same way.  Each column's slice is padded to an odd stride S so that lane c's row r,
// at c*S + r, falls in a different bank for every lane.  The arithmetic is unchanged.
constexpr int kColBlock = 64;
constexpr int kMaxStagedRows = 127;
// dynamic LDS one block may use: the device's per-block limit, read once
size_t max_block_lds() {
  static size_t cap = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
      return static_cast<size_t>(64 * 1024);
    return static_cast<size_t>(v);
  }();
  return cap;
}

__host__ __device__ inline int staged_stride(int n) { return n | 1; }

// Each lane loads the base dof of one column and the offsets of rows lane and lane + 64 once;
// the bases are then broadcast by lane shuffles, and the loads of kStageBatch columns are all
// issued before any of them is written to LDS, so a block waits for memory 64 / kStageBatch
// times.  n <= 2 * kColBlock.
constexpr int kStageBatch = 16;

template <typename T, int B = kStageBatch>
__device__ inline void stage_columns_in(T *__restrict__ xs, int S, const T *__restrict__ x,
                                        const int *__restrict__ map, int ndf, const int *__restrict__ ind,
                                        int n, int pos0, int nc, int cfirst = 0, int clast = kColBlock) {
  const int lane = threadIdx.x & (kColBlock - 1);
  const int mybase = pos0 + lane < nc ? map[static_cast<size_t>(pos0 + lane) * ndf] : 0;
  const int r0 = lane, r1 = lane + kColBlock;
  const int o0 = r0 < n ? ind[r0] : 0, o1 = r1 < n ? ind[r1] : 0;
  const int ncol = min(kColBlock, nc - pos0);
  for (int c0 = cfirst; c0 < clast; c0 += B) {
    T v0[B], v1[B];
#pragma unroll
    for (int u = 0; u < B; ++u) {
      const int base = __shfl(mybase, c0 + u, kColBlock);
      const bool ok = c0 + u < ncol;
      v0[u] = ok && r0 < n ? x[static_cast<size_t>(base) + o0] : T(0);
      v1[u] = ok && r1 < n ? x[static_cast<size_t>(base) + o1] : T(0);
    }
#pragma unroll
    for (int u = 0; u < B; ++u) {
      if (r0 < n) xs[(c0 + u) * S + r0] = v0[u];
      if (r1 < n) xs[(c0 + u) * S + r1] = v1[u];
    }
  }
}

template <typename T>
__device__ inline void stage_columns_out(const T *__restrict__ xs, int S, T *__restrict__ y,
                                         const int *__restrict__ map, int ndf, const int *__restrict__ ind,
                                         int n, int pos0, int nc, int cfirst = 0, int clast = kColBlock) {
  const int lane = threadIdx.x & (kColBlock - 1);
  const int mybase = pos0 + lane < nc ? map[static_cast<size_t>(pos0 + lane) * ndf] : 0;
  const int r0 = lane, r1 = lane + kColBlock;
  const int o0 = r0 < n ? ind[r0] : 0, o1 = r1 < n ? ind[r1] : 0;
  const int ncol = min(kColBlock, nc - pos0);
#pragma unroll 16
  for (int c = cfirst; c < clast; ++c) {
    const int base = __shfl(mybase, c, kColBlock);
    if (c < ncol) {
      if (r0 < n) y[static_cast<size_t>(base) + o0] = xs[c * S + r0];
      if (r1 < n) y[static_cast<size_t>(base) + o1] = xs[c * S + r1];
    }
  }
}

// k_cma_appinv with x and d' in LDS; c' in LDS too when CW_LDS (else in the coalesced work
// array).  The forward sweep loads CH rows of bands at a time, so 3*CH loads are in flight per
// lane.
template <typename T, int CH, bool CW_LDS>
__global__ __launch_bounds__(kColBlock) void k_cma_appinv_lds(const T *__restrict__ C, CmaP p, int nc,
                                                              const T *__restrict__ x, T *__restrict__ y,
                                                              const int *__restrict__ map, int ndf,
                                                              const int *__restrict__ ind, T *__restrict__ cw) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *xs = reinterpret_cast<T *>(lds_raw);
  const int nrow = p.nrow, S = staged_stride(nrow);
  T *cs = xs + kColBlock * S;
  const int pos0 = blockIdx.x * kColBlock;
  const int pos = pos0 + threadIdx.x;
  // the first chunk of bands is requested before the staging, so the two overlap
  T a1[CH], a2[CH], a3[CH];
#pragma unroll
  for (int u = 0; u < CH; ++u) {
    const int i = 1 + u;
    if (i > nrow || pos >= nc) continue;
    a1[u] = C[cidx(p, nc, i, 1, pos)];
    a2[u] = C[cidx(p, nc, i, 2, pos)];
    a3[u] = C[cidx(p, nc, i, 3, pos)];
  }
  stage_columns_in<T, sizeof(T) == 4 ? 32 : kStageBatch>(xs, S, x, map, ndf, ind, nrow, pos0, nc);
  __syncthreads();
  if (pos < nc) {
    T *xc = xs + threadIdx.x * S, *cc = cs + threadIdx.x * S;
    auto cw_at = [&](int i) -> T & {
      if constexpr (CW_LDS) return cc[i - 1];
      else return cw[static_cast<size_t>(i - 1) * nc + pos];
    };
    T cprev = T(0), dprev = T(0);
    for (int i0 = 1; i0 <= nrow; i0 += CH) {
      if (i0 > 1) {
#pragma unroll
        for (int u = 0; u < CH; ++u) {
          const int i = i0 + u;
          if (i > nrow) continue;
          a1[u] = C[cidx(p, nc, i, 1, pos)];
          a2[u] = C[cidx(p, nc, i, 2, pos)];
          a3[u] = C[cidx(p, nc, i, 3, pos)];
        }
      }
      // the chunk's right-hand sides are read from LDS before any of its results are written,
      // so the recurrence does not wait on an LDS round trip per row
      T xv[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int i = i0 + u;
        if (i > nrow) continue;
        xv[u] = xc[i - 1];
      }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int i = i0 + u;
        if (i > nrow) continue;
        T inv, cp, dp;
        if (i == 1) {
          inv = T(1) / a2[u];
          cp = inv * a3[u];
          dp = inv * xv[u];
        } else {
          inv = T(1) / (a2[u] - a1[u] * cprev);
          cp = i < nrow ? inv * a3[u] : T(0);
          dp = inv * (xv[u] - a1[u] * dprev);
        }
        cw_at(i) = cp;
        xc[i - 1] = dp;
        cprev = cp;
        dprev = dp;
      }
    }
    T dnext = dprev;
    for (int i0 = nrow - 1; i0 >= 1; i0 -= CH) {
      T cv[CH], dv[CH];
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int i = i0 - u;
        if (i < 1) continue;
        cv[u] = cw_at(i);
        dv[u] = xc[i - 1];
      }
#pragma unroll
      for (int u = 0; u < CH; ++u) {
        const int i = i0 - u;
        if (i < 1) continue;
        const T d = dv[u] - cv[u] * dnext;
        xc[i - 1] = d;
        dnext = d;
      }
    }
  }
  __syncthreads();
  stage_columns_out(xs, S, y, map, ndf, ind, nrow, pos0, nc);
}

// k_cma_apply<T, BW> for 64 columns per block of kApplyWaves wavefronts: the waves stage x
// (and y when accumulating) for a quarter of the columns each, then every wave computes a
// contiguous quarter of the rows of all 64 columns (lane = column), and the waves store y back a
// quarter of the columns each.  Rows are independent in an apply, so this splits a column's
// work four ways; a block needs kColBlock*(Sx + Sy) values of LDS.
constexpr int kApplyWaves = 4;

// A wave's rows of bands are loaded kApplyChunk at a time; the first chunk is issued before
// the fields are staged, so the two sets of loads are in flight together.
constexpr int kApplyChunk = 18;

template <typename T, int BW>
__device__ inline void load_band_chunk(T (&a)[kApplyChunk][BW], const T *__restrict__ C, const CmaP &p, int nc,
                                       int pos, int i0, int ilast) {
#pragma unroll
  for (int u = 0; u < kApplyChunk; ++u) {
    const int i = i0 + u;
    if (i > ilast) continue;
    const int jm = jminus(p, i), j0 = max(1, jm), j1 = min(p.ncol, jplus(p, i));
#pragma unroll
    for (int b = 0; b < BW; ++b) {
      const int j = jm + b;
      a[u][b] = (j >= j0 && j <= j1) ? C[cidx(p, nc, i, b + 1, pos)] : T(0);
    }
  }
}

// four blocks of FP32 staging fit a CU's LDS; at most 128 VGPRs keeps all four resident
template <typename T, int BW>
__global__ __launch_bounds__(kColBlock * kApplyWaves) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2)))
void k_cma_apply_lds(
    const T *__restrict__ C, CmaP p, int nc, const T *__restrict__ x, T *__restrict__ y,
    const int *__restrict__ map_to, int ndf_to, const int *__restrict__ map_from, int ndf_from,
    const int *__restrict__ ind_to, const int *__restrict__ ind_from, int accumulate) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T *xs = reinterpret_cast<T *>(lds_raw);
  const int Sx = staged_stride(p.ncol), Sy = staged_stride(p.nrow);
  T *ys = xs + kColBlock * Sx;
  const int pos0 = blockIdx.x * kColBlock;
  const int w = threadIdx.x / kColBlock, lane = threadIdx.x & (kColBlock - 1);
  constexpr int kCw = kColBlock / kApplyWaves;
  const int pos = pos0 + lane;
  const int q = (p.nrow + kApplyWaves - 1) / kApplyWaves;
  const int ifirst = 1 + w * q, ilast = min(p.nrow, (w + 1) * q);
  T a[kApplyChunk][BW];
  if (pos < nc) load_band_chunk<T, BW>(a, C, p, nc, pos, ifirst, ilast);
  stage_columns_in(xs, Sx, x, map_from, ndf_from, ind_from, p.ncol, pos0, nc, w * kCw, (w + 1) * kCw);
  if (accumulate) stage_columns_in(ys, Sy, y, map_to, ndf_to, ind_to, p.nrow, pos0, nc, w * kCw, (w + 1) * kCw);
  __syncthreads();
  if (pos < nc) {
    const T *xc = xs + lane * Sx;
    T *yc = ys + lane * Sy;
    for (int i0 = ifirst; i0 <= ilast; i0 += kApplyChunk) {
      if (i0 != ifirst) load_band_chunk<T, BW>(a, C, p, nc, pos, i0, ilast);
#pragma unroll
      for (int u = 0; u < kApplyChunk; ++u) {
        const int i = i0 + u;
        if (i > ilast) continue;
        const int jm = jminus(p, i), j0 = max(1, jm), j1 = min(p.ncol, jplus(p, i));
        T acc = accumulate ? yc[i - 1] : T(0);
#pragma unroll
        for (int b = 0; b < BW; ++b) {
          const int j = jm + b;
          if (j >= j0 && j <= j1) acc = acc + a[u][b] * xc[j - 1];
        }
        yc[i - 1] = acc;
      }
    }
  }
  __syncthreads();
  stage_columns_out(ys, Sy, y, map_to, ndf_to, ind_to, p.nrow, pos0, nc, w * kCw, (w + 1) * kCw);
}

// columnwise_op_mul_kernel_code (sci_columnwise_op_mul_kernel_mod.F90:128-160): C += A B.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cma_mul(const T *__restrict__ A, CmaP pa, const T *__restrict__ B, CmaP pb,
                                                    T *__restrict__ C, CmaP pc, int nc) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos >= nc) return;
  for (int i = 1; i <= pc.nrow; ++i) {
    const int jma = jminus(pa, i), jpa = jplus(pa, i), jmc = jminus(pc, i), jpc = jplus(pc, i);
    for (int j = max(1, jmc); j <= min(pc.ncol, jpc); ++j) {
      T e = T(0);
      for (int l = max(1, jma); l <= min(pa.ncol, jpa); ++l) {
        const int jmb = jminus(pb, l), jpb = jplus(pb, l);
        if (jmb <= j && j <= jpb) e = e + A[cidx(pa, nc, i, l - jma + 1, pos)] * B[cidx(pb, nc, l, j - jmb + 1, pos)];
      }
      const size_t c = cidx(pc, nc, i, j - jmc + 1, pos);
      C[c] = C[c] + e;
    }
  }
}

// columnwise_op_scaledadd_kernel_code (sci_columnwise_op_scaledadd_kernel_mod.F90:132-166):
// C = alpha A + beta B, each added band by band onto a zeroed C (B's loop bounded by A's ncol,
// as there).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cma_scaledadd(const T *__restrict__ A, CmaP pa, const T *__restrict__ B,
                                                          CmaP pb, T *__restrict__ C, CmaP pc, int nc, T alpha,
                                                          T beta) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos >= nc) return;
  for (int i = 1; i <= pc.nrow; ++i)
    for (int b = 1; b <= pc.bw; ++b) C[cidx(pc, nc, i, b, pos)] = T(0);
  for (int i = 1; i <= pa.nrow; ++i) {
    const int jma = jminus(pa, i), jpa = jplus(pa, i), jmb = jminus(pb, i), jpb = jplus(pb, i), jmc = jminus(pc, i);
    for (int j = max(1, jma); j <= min(pa.ncol, jpa); ++j) {
      const size_t c = cidx(pc, nc, i, j - jmc + 1, pos);
      C[c] = C[c] + alpha * A[cidx(pa, nc, i, j - jma + 1, pos)];
    }
    for (int j = max(1, jmb); j <= min(pa.ncol, jpb); ++j) {
      const size_t c = cidx(pc, nc, i, j - jmc + 1, pos);
      C[c] = C[c] + beta * B[cidx(pb, nc, i, j - jmb + 1, pos)];
    }
  }
}

// columnwise_op_asm_diag_hmht_kernel_code (sci_columnwise_op_asm_diag_hmht_kernel_mod.F90:
// 95-135): the W3 column band of Dh diag(M2h)^-1 Dh^T, tmp = sum over df3 of
// Dh(df1,df3)*Dh(df2,df3)/M2h(df3,df3).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cma_asm_hmht(T *__restrict__ C, CmaP p, int nc, const T *__restrict__ Dh,
                                                         size_t gsd, const int *__restrict__ ebd,
                                                         const T *__restrict__ M, size_t gsm,
                                                         const int *__restrict__ ebm, int nl, int ndf3, int ndf2h,
                                                         const int *__restrict__ band) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos >= nc) return;
  for (int i = 1; i <= p.nrow; ++i)
    for (int b = 1; b <= p.bw; ++b) C[cidx(p, nc, i, b, pos)] = T(0);
  for (int k = 0; k < nl; ++k) {
    const long long ed = ebd[pos] + k, em = ebm[pos] + k;
    for (int df1 = 0; df1 < ndf3; ++df1) {
      const int i = band[k * ndf3 + df1];
      const int jm = jminus(p, i);
      for (int df2 = 0; df2 < ndf3; ++df2) {
        T tmp = T(0);
        for (int df3 = 0; df3 < ndf2h; ++df3)
          tmp = tmp + Dh[aidx<T>(gsd, ed, df1 * ndf2h + df3)] * Dh[aidx<T>(gsd, ed, df2 * ndf2h + df3)] /
                          M[aidx<T>(gsm, em, df3 * ndf2h + df3)];
        const int j = band[k * ndf3 + df2];
        const size_t c = cidx(p, nc, i, j - jm + 1, pos);
        C[c] = C[c] + tmp;
      }
    }
  }
}

// columnwise_op_asm_m2v_lumped_inv_krnl_code (sci_columnwise_op_asm_m2v_lumped_inv_krnl_mod.F90:
// 85-113): band 1 of each row sums the diagonal entries that land on it, then every nonzero
// entry of the column becomes its reciprocal.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cma_asm_lumped_inv(T *__restrict__ C, CmaP p, int nc,
                                                               const T *__restrict__ A, size_t gs,
                                                               const int *__restrict__ ebase, uint32_t row_top,
                                                               int nl, int ndf, const int *__restrict__ band) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos >= nc) return;
  for (int i = 1; i <= p.nrow; ++i)
    for (int b = 1; b <= p.bw; ++b) C[cidx(p, nc, i, b, pos)] = T(0);
  for (int k = 0; k < nl; ++k) {
    const long long el = ebase[pos] + k;
    for (int df = 0; df < ndf; ++df) {
      const size_t c = cidx(p, nc, band[k * ndf + df], 1, pos);
      C[c] = C[c] + A[aidx<T>(gs, el + ((row_top >> df) & 1u), df * ndf + df)];
    }
  }
  for (int i = 1; i <= p.nrow; ++i)
    for (int b = 1; b <= p.bw; ++b) {
      const size_t c = cidx(p, nc, i, b, pos);
      const T v = C[c];
      if (v != T(0)) C[c] = T(1) / v;
    }
}

// tri_solve_code (components/science/source/kernel/solver/sci_tri_solve_kernel_mod.f90:57-130):
// W3 columns of nl levels, diagonal t0, super-diagonal tp, sub-diagonal tm; tp' is kept in y
// during the forward sweep.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tri_solve(const T *__restrict__ x, const T *__restrict__ t0,
                                                      const T *__restrict__ tp, const T *__restrict__ tm,
                                                      T *__restrict__ y, T *__restrict__ xw,
                                                      const int *__restrict__ map, int nc, int nl) {
  const int pos = blockIdx.x * kBlock + threadIdx.x;
  if (pos >= nc) return;
  const size_t ij = static_cast<size_t>(map[pos]);
  T denom = T(1) / t0[ij];
  T tpn = tp[ij] * denom, xn = x[ij] * denom;
  y[ij] = tpn;
  xw[ij] = xn;
  for (int k = 1; k <= nl - 1; ++k) {
    denom = T(1) / (t0[ij + k] - tm[ij + k] * tpn);
    tpn = tp[ij + k] * denom;
    xn = (x[ij + k] - tm[ij + k] * xn) * denom;
    y[ij + k] = tpn;
    xw[ij + k] = xn;
  }
  T ynext = xn;
  y[ij + nl - 1] = ynext;
  for (int k = nl - 2; k >= 0; --k) {
    ynext = xw[ij + k] - y[ij + k] * ynext;
    y[ij + k] = ynext;
  }
}

#pragma clang fp contract(on)

// reference layout (bw, nrow, ncell_2d) <-> device layout, by cell of each position
template <typename T>
__global__ void k_cma_from_ref(T *__restrict__ C, CmaP p, int nc, const double *__restrict__ R,
                               const int *__restrict__ cell_of_pos) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long per = static_cast<long long>(p.bw) * p.nrow;
  if (t >= per * nc) return;
  const int pos = static_cast<int>(t / per);
  const int r = static_cast<int>(t - static_cast<long long>(pos) * per);
  const int b = r % p.bw, i = r / p.bw;
  C[(static_cast<size_t>(i) * p.bw + b) * nc + pos] = static_cast<T>(R[static_cast<size_t>(cell_of_pos[pos]) * per + r]);
}

template <typename T>
__global__ void k_cma_to_ref(const T *__restrict__ C, CmaP p, int nc, double *__restrict__ R,
                             const int *__restrict__ cell_of_pos) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long per = static_cast<long long>(p.bw) * p.nrow;
  if (t >= per * nc) return;
  const int pos = static_cast<int>(t / per);
  const int r = static_cast<int>(t - static_cast<long long>(pos) * per);
  const int b = r % p.bw, i = r / p.bw;
  R[static_cast<size_t>(cell_of_pos[pos]) * per + r] = static_cast<double>(C[(static_cast<size_t>(i) * p.bw + b) * nc + pos]);
}

}  // namespace

#define CLAUNCH(kernel, n, ...) hipLaunchKernelGGL((kernel), dim3(grid_for(n)), dim3(kBlock), 0, cur_stream(), __VA_ARGS__)

// Interior / vertical-face dof counts of a horizontally discontinuous lowest-order space, from
// its column roles: nf bottom faces first, then ni interior dofs, then the nf top faces (the
// order columnwise_operator_mod's banded dofmap assumes), every column owned by one cell and
// every cell laid out like the first.
static int cma_space(const FunctionSpace &f, int &ni, int &nf) {
  nf = __builtin_popcount(f.top_mask);
  ni = f.ndf - 2 * nf;
  LF_CHECK(ni >= 0, LFRIC_E_UNSUPPORTED, "columnwise operator: space roles not bottom/interior/top");
  for (int j = 0; j < nf; ++j)
    LF_CHECK(f.partner[j] == nf + ni + j, LFRIC_E_UNSUPPORTED,
             "columnwise operator: dofs must be ordered bottom faces, interior, top faces");
  std::vector<int> owner(static_cast<size_t>(f.undf) + 1, -1);
  for (int c = 0; c < f.ncells; ++c)
    for (int d = 0; d < f.ndf; ++d) {
      const int m = f.dofmap[static_cast<size_t>(c) * f.ndf + d];
      LF_CHECK(m - f.dofmap[static_cast<size_t>(c) * f.ndf] == f.dofmap[d] - f.dofmap[0], LFRIC_E_UNSUPPORTED,
               "columnwise operator: every cell's column must be laid out like the first");
      if ((f.top_mask >> d) & 1u) continue;
      LF_CHECK(owner[m] == -1 || owner[m] == c, LFRIC_E_UNSUPPORTED,
               "columnwise operator: the space is not horizontally discontinuous");
      owner[m] = c;
    }
  return LFRIC_OK;
}

static int upload_ints(const std::vector<int> &v, int **d, size_t &bytes) {
  bytes = sizeof(int) * v.size();
  if (int rc = dmalloc(reinterpret_cast<void **>(d), bytes)) return rc;
  LF_HIP(hipMemcpy(*d, v.data(), bytes, hipMemcpyHostToDevice));
  return LFRIC_OK;
}

// build_dofmaps (columnwise_operator_mod.f90:196-230): banded row of (df, k) and the column
// offset (map(df) - map(1) + k) of every banded row, from the first cell.
static void cma_dofmaps(const FunctionSpace &f, int ni, int nf, std::vector<int> &band, std::vector<int> &ind) {
  const int nl = f.nlayers, ndf = f.ndf;
  band.assign(static_cast<size_t>(nl) * ndf, 0);
  ind.assign(static_cast<size_t>(nl) * ni + static_cast<size_t>(nl + 1) * nf, 0);
  for (int k = 0; k < nl; ++k)
    for (int df = 0; df < ndf; ++df) {
      const int b = k * (ni + nf) + df + 1;
      band[static_cast<size_t>(k) * ndf + df] = b;
      ind[b - 1] = f.dofmap[df] - f.dofmap[0] + k;
    }
}

static size_t cma_esize(int prec) { return prec == LFRIC_R64 ? sizeof(double) : sizeof(float); }

static int cma_create(const lfric_fs *to, const lfric_fs *from, int precision, bool custom, int alpha, int beta,
                      int gamma_m, int gamma_p, lfric_cma **out) {
  if (int rc = require_init()) return rc;
  LF_CHECK(to && from && out, LFRIC_E_ARG, "null argument");
  LF_CHECK(precision == LFRIC_R32 || precision == LFRIC_R64, LFRIC_E_ARG, "columnwise operator: R32 or R64");
  const FunctionSpace &ft = *to->fs, &ff = *from->fs;
  LF_CHECK(ft.nlayers == ff.nlayers && ft.ncells == ff.ncells && ft.cell_of_pos == ff.cell_of_pos, LFRIC_E_ARG,
           "columnwise operator: spaces on different meshes or colourings");
  auto A = std::make_unique<lfric_cma>();
  A->to = to->fs;
  A->from = from->fs;
  A->precision = precision;
  if (int rc = cma_space(ft, A->ni_to, A->nf_to)) return rc;
  if (int rc = cma_space(ff, A->ni_from, A->nf_from)) return rc;
  // columnwise_operator_initialiser(_custom), divide_by_gcd (columnwise_operator_mod.f90:89-190)
  if (!custom) {
    alpha = A->ni_from + A->nf_from;
    beta = A->ni_to + A->nf_to;
    gamma_m = (ff.ndf - 1) * (A->ni_to + A->nf_to);
    gamma_p = (ft.ndf - 1) * (A->ni_from + A->nf_from);
  }
  LF_CHECK(alpha > 0 && beta > 0 && gamma_m >= 0 && gamma_p >= 0, LFRIC_E_ARG,
           "columnwise operator: alpha, beta > 0 and gamma >= 0 expected");
  const int g = std::gcd(alpha, beta);
  if (g > 1) {
    alpha /= g;
    beta /= g;
    gamma_m /= g;
    gamma_p /= g;
  }
  A->alpha = alpha;
  A->beta = beta;
  A->gamma_m = gamma_m;
  A->gamma_p = gamma_p;
  A->bw = 1 + (gamma_m + gamma_p + beta - 1) / beta;
  A->nrow = ft.nlayers * A->ni_to + (ft.nlayers + 1) * A->nf_to;
  A->ncol = ff.nlayers * A->ni_from + (ff.nlayers + 1) * A->nf_from;
  A->ncols = ft.ncells;
  A->bytes = cma_esize(precision) * static_cast<size_t>(A->bw) * A->nrow * A->ncols;
  if (int rc = dmalloc(&A->d, A->bytes)) return rc;
  LF_HIP(hipMemset(A->d, 0, A->bytes));
  std::vector<int> band, ind;
  cma_dofmaps(ft, A->ni_to, A->nf_to, band, ind);
  if (int rc = upload_ints(band, &A->d_band_to, A->band_to_bytes)) return rc;
  if (int rc = upload_ints(ind, &A->d_ind_to, A->ind_to_bytes)) return rc;
  cma_dofmaps(ff, A->ni_from, A->nf_from, band, ind);
  if (int rc = upload_ints(band, &A->d_band_from, A->band_from_bytes)) return rc;
  if (int rc = upload_ints(ind, &A->d_ind_from, A->ind_from_bytes)) return rc;
  *out = A.release();
  return LFRIC_OK;
}

// The band each (row, column) pair of an assembly lands in must lie inside the operator's
// bandwidth (an operator created with too narrow custom parameters would be written out of
// bounds): checked on the host before any assembly kernel runs.
static int check_assembly_fits(const lfric_cma &C, bool diagonal_pairs_only) {
  std::vector<int> bt, it, bf, itf;
  cma_dofmaps(*C.to, C.ni_to, C.nf_to, bt, it);
  cma_dofmaps(*C.from, C.ni_from, C.nf_from, bf, itf);
  const CmaP p = params_of(C);
  const int nl = C.to->nlayers, n1 = C.to->ndf, n2 = C.from->ndf;
  for (int k = 0; k < nl; ++k)
    for (int a = 0; a < n1; ++a)
      for (int b = 0; b < n2; ++b) {
        if (diagonal_pairs_only && a != b) continue;
        const int i = bt[static_cast<size_t>(k) * n1 + a];
        const int j = (diagonal_pairs_only ? bt : bf)[static_cast<size_t>(k) * (diagonal_pairs_only ? n1 : n2) + b];
        const int band = j - jminus(p, i) + 1;
        LF_CHECK(i >= 1 && i <= p.nrow && band >= 1 && band <= p.bw, LFRIC_E_ARG,
                 "columnwise assembly: an entry falls outside the operator's band (custom parameters too narrow)");
      }
  return LFRIC_OK;
}

static int finish(const char *what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(LFRIC_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return LFRIC_OK;
}

int launch_cma_appinv(const lfric_cma *A, const void *x, void *y) {
  LF_CHECK(A->bw == 3 && A->nrow == A->ncol && A->to == A->from, LFRIC_E_UNSUPPORTED,
           "columnwise inverse: a square tridiagonal operator is needed");
  const size_t wb = cma_esize(A->precision) * static_cast<size_t>(A->nrow) * A->ncols;
  if (A->work_bytes < wb) {
    dfree(A->work, A->work_bytes);
    A->work_bytes = 0;
    if (int rc = dmalloc(&A->work, wb)) return rc;
    A->work_bytes = wb;
  }
  const CmaP p = params_of(*A);
  const FunctionSpace &f = *A->to;
  const size_t lds1 = cma_esize(A->precision) * kColBlock * staged_stride(A->nrow);
  if (A->nrow <= kMaxStagedRows && lds1 <= max_block_lds()) {
    const int blocks = (A->ncols + kColBlock - 1) / kColBlock;
    const bool cw_lds = 2 * lds1 <= 64 * 1024;  // c' beside x while four blocks still fit a CU
    const size_t lds = cw_lds ? 2 * lds1 : lds1;
#define APPINV_LDS(T, CH)                                                                                       \
  if (cw_lds)                                                                                                   \
    hipLaunchKernelGGL((k_cma_appinv_lds<T, CH, true>), dim3(blocks), dim3(kColBlock), lds, cur_stream(),       \
                       static_cast<const T *>(A->d), p, A->ncols, static_cast<const T *>(x), static_cast<T *>(y), \
                       f.d_map, f.ndf, A->d_ind_to, static_cast<T *>(A->work));                                 \
  else                                                                                                          \
    hipLaunchKernelGGL((k_cma_appinv_lds<T, CH, false>), dim3(blocks), dim3(kColBlock), lds, cur_stream(),      \
                       static_cast<const T *>(A->d), p, A->ncols, static_cast<const T *>(x), static_cast<T *>(y), \
                       f.d_map, f.ndf, A->d_ind_to, static_cast<T *>(A->work));
    if (A->precision == LFRIC_R64) {
      APPINV_LDS(double, 36)
    } else {
      APPINV_LDS(float, 72)
    }
#undef APPINV_LDS
    return finish("columnwise inverse");
  }
  if (A->precision == LFRIC_R64)
    CLAUNCH(k_cma_appinv<double>, A->ncols, static_cast<const double *>(A->d), p, A->ncols,
            static_cast<const double *>(x), static_cast<double *>(y), f.d_map, f.ndf, A->d_ind_to,
            static_cast<double *>(A->work));
  else
    CLAUNCH(k_cma_appinv<float>, A->ncols, static_cast<const float *>(A->d), p, A->ncols,
            static_cast<const float *>(x), static_cast<float *>(y), f.d_map, f.ndf, A->d_ind_to,
            static_cast<float *>(A->work));
  return finish("columnwise inverse");
}

}  // namespace lfric

using namespace lfric;

lfric_cma::~lfric_cma() {
  dfree(d, bytes);
  dfree(d_ind_to, ind_to_bytes);
  dfree(d_ind_from, ind_from_bytes);
  dfree(d_band_to, band_to_bytes);
  dfree(d_band_from, band_from_bytes);
  dfree(work, work_bytes);
}

extern "C" {

int lfric_cma_create(const lfric_fs *to, const lfric_fs *from, int precision, lfric_cma **A) {
  return cma_create(to, from, precision, false, 0, 0, 0, 0, A);
}

int lfric_cma_create_custom(const lfric_fs *to, const lfric_fs *from, int precision, int alpha, int beta, int gamma_m,
                            int gamma_p, lfric_cma **A) {
  return cma_create(to, from, precision, true, alpha, beta, gamma_m, gamma_p, A);
}

int lfric_cma_info(const lfric_cma *A, int *params) {
  LF_CHECK(A && params, LFRIC_E_ARG, "null argument");
  const int v[7] = {A->alpha, A->beta, A->gamma_m, A->gamma_p, A->bw, A->nrow, A->ncol};
  std::copy(v, v + 7, params);
  return LFRIC_OK;
}

void lfric_cma_destroy(lfric_cma *A) { delete A; }

static int cma_ref_copy(const lfric_cma *A, const double *in, double *out) {
  if (int rc = require_init()) return rc;
  const FunctionSpace &f = *A->to;
  const size_t n = static_cast<size_t>(A->bw) * A->nrow * A->ncols;
  double *d_ref = nullptr;
  int *d_cop = nullptr;
  if (int rc = dmalloc(reinterpret_cast<void **>(&d_ref), sizeof(double) * n)) return rc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&d_cop), sizeof(int) * f.cell_of_pos.size())) {
    dfree(d_ref, sizeof(double) * n);
    return rc;
  }
  hipError_t e = hipMemcpy(d_cop, f.cell_of_pos.data(), sizeof(int) * f.cell_of_pos.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess && in) e = hipMemcpy(d_ref, in, sizeof(double) * n, hipMemcpyHostToDevice);
  const CmaP p = params_of(*A);
  if (e == hipSuccess) {
    if (in) {
      if (A->precision == LFRIC_R64)
        CLAUNCH(k_cma_from_ref<double>, static_cast<long long>(n), static_cast<double *>(A->d), p, A->ncols, d_ref, d_cop);
      else
        CLAUNCH(k_cma_from_ref<float>, static_cast<long long>(n), static_cast<float *>(A->d), p, A->ncols, d_ref, d_cop);
    } else {
      if (A->precision == LFRIC_R64)
        CLAUNCH(k_cma_to_ref<double>, static_cast<long long>(n), static_cast<const double *>(A->d), p, A->ncols, d_ref, d_cop);
      else
        CLAUNCH(k_cma_to_ref<float>, static_cast<long long>(n), static_cast<const float *>(A->d), p, A->ncols, d_ref, d_cop);
    }
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(cur_stream());
  if (e == hipSuccess && out) e = hipMemcpy(out, d_ref, sizeof(double) * n, hipMemcpyDeviceToHost);
  dfree(d_ref, sizeof(double) * n);
  dfree(d_cop, sizeof(int) * f.cell_of_pos.size());
  if (e != hipSuccess) return fail(LFRIC_E_HIP, hipGetErrorString(e));
  return LFRIC_OK;
}

int lfric_cma_upload(lfric_cma *A, const double *ref) {
  LF_CHECK(A && ref, LFRIC_E_ARG, "null argument");
  return cma_ref_copy(A, ref, nullptr);
}

int lfric_cma_download(const lfric_cma *A, double *ref) {
  LF_CHECK(A && ref, LFRIC_E_ARG, "null argument");
  return cma_ref_copy(A, nullptr, ref);
}

int lfric_cma_assemble(lfric_cma *C, const lfric_op *A) {
  if (int rc = require_init()) return rc;
  LF_CHECK(C && A, LFRIC_E_ARG, "null argument");
  LF_CHECK(A->to == C->to && A->from == C->from && A->precision == C->precision, LFRIC_E_ARG,
           "columnwise assembly: the LMA operator must have the same spaces and precision");
  if (int rc = check_assembly_fits(*C, false)) return rc;
  const CmaP p = params_of(*C);
  const FunctionSpace &f = *C->to;
  if (C->precision == LFRIC_R64)
    CLAUNCH(k_cma_asm<double>, C->ncols, static_cast<double *>(C->d), p, C->ncols, static_cast<const double *>(A->d),
            A->gs, f.d_ebase, A->row_top, f.nlayers, f.ndf, C->from->ndf, C->d_band_to, C->d_band_from);
  else
    CLAUNCH(k_cma_asm<float>, C->ncols, static_cast<float *>(C->d), p, C->ncols, static_cast<const float *>(A->d),
            A->gs, f.d_ebase, A->row_top, f.nlayers, f.ndf, C->from->ndf, C->d_band_to, C->d_band_from);
  return finish("columnwise assembly");
}

int lfric_cma_assemble_diag_hmht(lfric_cma *C, const lfric_op *Dh, const lfric_op *M2h) {
  if (int rc = require_init()) return rc;
  LF_CHECK(C && Dh && M2h, LFRIC_E_ARG, "null argument");
  LF_CHECK(C->to == C->from && Dh->to == C->to && M2h->to == Dh->from && M2h->from == Dh->from, LFRIC_E_ARG,
           "diag_hmht: Dh maps W2h to the operator's space and M2h is square on W2h");
  LF_CHECK(Dh->precision == C->precision && M2h->precision == C->precision, LFRIC_E_ARG, "diag_hmht: precisions differ");
  LF_CHECK(M2h->row_top == 0u && Dh->row_top == 0u, LFRIC_E_UNSUPPORTED, "diag_hmht: W2h has no vertical faces");
  if (int rc = check_assembly_fits(*C, false)) return rc;
  const CmaP p = params_of(*C);
  const FunctionSpace &f = *C->to;
  if (C->precision == LFRIC_R64)
    CLAUNCH(k_cma_asm_hmht<double>, C->ncols, static_cast<double *>(C->d), p, C->ncols,
            static_cast<const double *>(Dh->d), Dh->gs, Dh->to->d_ebase, static_cast<const double *>(M2h->d), M2h->gs,
            M2h->to->d_ebase, f.nlayers, f.ndf, Dh->from->ndf, C->d_band_to);
  else
    CLAUNCH(k_cma_asm_hmht<float>, C->ncols, static_cast<float *>(C->d), p, C->ncols,
            static_cast<const float *>(Dh->d), Dh->gs, Dh->to->d_ebase, static_cast<const float *>(M2h->d), M2h->gs,
            M2h->to->d_ebase, f.nlayers, f.ndf, Dh->from->ndf, C->d_band_to);
  return finish("diag_hmht");
}

int lfric_cma_assemble_m2v_lumped_inv(lfric_cma *C, const lfric_op *M) {
  if (int rc = require_init()) return rc;
  LF_CHECK(C && M, LFRIC_E_ARG, "null argument");
  LF_CHECK(M->to == C->to && M->from == C->to && C->to == C->from && M->precision == C->precision, LFRIC_E_ARG,
           "lumped inverse: a square LMA operator on the operator's space and precision is needed");
  const CmaP p = params_of(*C);
  const FunctionSpace &f = *C->to;
  if (C->precision == LFRIC_R64)
    CLAUNCH(k_cma_asm_lumped_inv<double>, C->ncols, static_cast<double *>(C->d), p, C->ncols,
            static_cast<const double *>(M->d), M->gs, f.d_ebase, M->row_top, f.nlayers, f.ndf, C->d_band_to);
  else
    CLAUNCH(k_cma_asm_lumped_inv<float>, C->ncols, static_cast<float *>(C->d), p, C->ncols,
            static_cast<const float *>(M->d), M->gs, f.d_ebase, M->row_top, f.nlayers, f.ndf, C->d_band_to);
  return finish("lumped inverse");
}

static int check_cma_vecs(const lfric_cma *A, const lfric_vec *x, const lfric_vec *y) {
  LF_CHECK(A && x && y, LFRIC_E_ARG, "null argument");
  LF_CHECK(x->fs->undf == A->from->undf && y->fs->undf == A->to->undf && x->nrhs == 1 && y->nrhs == 1, LFRIC_E_ARG,
           "columnwise operator: vectors do not match its spaces");
  LF_CHECK(x->precision == A->precision && y->precision == A->precision, LFRIC_E_ARG,
           "columnwise operator: vector precision differs");
  return LFRIC_OK;
}

int lfric_cma_apply(const lfric_cma *A, const lfric_vec *x, lfric_vec *y, int accumulate) {
  if (int rc = require_init()) return rc;
  if (int rc = check_cma_vecs(A, x, y)) return rc;
  LF_CHECK(x->d != y->d, LFRIC_E_ARG, "columnwise apply: x and y alias");
  const CmaP p = params_of(*A);
  const FunctionSpace &ft = *A->to, &ff = *A->from;
  const int sblocks = (A->ncols + kColBlock - 1) / kColBlock;
  const size_t slds = cma_esize(A->precision) * kColBlock * (staged_stride(A->nrow) + staged_stride(A->ncol));
  const bool staged = A->nrow <= kMaxStagedRows && A->ncol <= kMaxStagedRows && slds <= max_block_lds();
#define CMA_APPLY(T, BW)                                                                                          \
  if (staged && BW > 0)                                                                                           \
    hipLaunchKernelGGL((k_cma_apply_lds<T, (BW > 0 ? BW : 1)>), dim3(sblocks), dim3(kColBlock * kApplyWaves), slds,   \
                       cur_stream(),                                                                              \
                       static_cast<const T *>(A->d), p, A->ncols, static_cast<const T *>(x->d),                   \
                       static_cast<T *>(y->d), ft.d_map, ft.ndf, ff.d_map, ff.ndf, A->d_ind_to, A->d_ind_from,    \
                       accumulate);                                                                               \
  else                                                                                                            \
    CLAUNCH((k_cma_apply<T, BW>), A->ncols, static_cast<const T *>(A->d), p, A->ncols, static_cast<const T *>(x->d), \
            static_cast<T *>(y->d), ft.d_map, ft.ndf, ff.d_map, ff.ndf, A->d_ind_to, A->d_ind_from, accumulate)
#define CMA_APPLY_T(T)      \
  switch (A->bw) {          \
    case 1:                 \
      CMA_APPLY(T, 1);      \
      break;                \
    case 2:                 \
      CMA_APPLY(T, 2);      \
      break;                \
    case 3:                 \
      CMA_APPLY(T, 3);      \
      break;                \
    default:                \
      CMA_APPLY(T, 0);      \
  }
  if (A->precision == LFRIC_R64) {
    CMA_APPLY_T(double)
  } else {
    CMA_APPLY_T(float)
  }
#undef CMA_APPLY_T
#undef CMA_APPLY
  return finish("columnwise apply");
}

int lfric_cma_apply_inverse(const lfric_cma *A, const lfric_vec *x, lfric_vec *y) {
  if (int rc = require_init()) return rc;
  if (int rc = check_cma_vecs(A, x, y)) return rc;
  return launch_cma_appinv(A, x->d, y->d);
}

int lfric_cma_multiply(const lfric_cma *A, const lfric_cma *B, lfric_cma *C) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A && B && C, LFRIC_E_ARG, "null argument");
  LF_CHECK(A->from == B->to && C->to == A->to && C->from == B->from, LFRIC_E_ARG,
           "columnwise product: C = A B needs A: Y <- X, B: X <- Z, C: Y <- Z");
  LF_CHECK(A->precision == C->precision && B->precision == C->precision, LFRIC_E_ARG, "columnwise product: precisions");
  LF_CHECK(A->nrow == C->nrow && B->ncol == C->ncol && A->ncol == B->nrow, LFRIC_E_ARG, "columnwise product: sizes");
  if (C->precision == LFRIC_R64)
    CLAUNCH(k_cma_mul<double>, C->ncols, static_cast<const double *>(A->d), params_of(*A),
            static_cast<const double *>(B->d), params_of(*B), static_cast<double *>(C->d), params_of(*C), C->ncols);
  else
    CLAUNCH(k_cma_mul<float>, C->ncols, static_cast<const float *>(A->d), params_of(*A),
            static_cast<const float *>(B->d), params_of(*B), static_cast<float *>(C->d), params_of(*C), C->ncols);
  return finish("columnwise product");
}

int lfric_cma_scaled_add(const lfric_cma *A, const lfric_cma *B, lfric_cma *C, double alpha, double beta) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A && B && C && C != A && C != B, LFRIC_E_ARG, "null or aliased operators");
  LF_CHECK(A->to == C->to && B->to == C->to && A->from == C->from && B->from == C->from, LFRIC_E_ARG,
           "columnwise sum: operators on different spaces");
  LF_CHECK(A->precision == C->precision && B->precision == C->precision, LFRIC_E_ARG, "columnwise sum: precisions");
  LF_CHECK(A->nrow == C->nrow && B->nrow == C->nrow && A->ncol == C->ncol && B->ncol == C->ncol, LFRIC_E_ARG,
           "columnwise sum: sizes");
  {  // every band entry of A and B must have a place in C's band
    const CmaP pa = params_of(*A), pb = params_of(*B), pc = params_of(*C);
    for (int i = 1; i <= pa.nrow; ++i) {
      const int jmc = jminus(pc, i);
      for (const CmaP *q : {&pa, &pb}) {
        const int lo = std::max(1, jminus(*q, i)), hi = std::min(pa.ncol, jplus(*q, i));
        LF_CHECK(lo > hi || (lo - jmc + 1 >= 1 && hi - jmc + 1 <= pc.bw), LFRIC_E_ARG,
                 "columnwise sum: C's band is narrower than A's or B's");
      }
    }
  }
  if (C->precision == LFRIC_R64)
    CLAUNCH(k_cma_scaledadd<double>, C->ncols, static_cast<const double *>(A->d), params_of(*A),
            static_cast<const double *>(B->d), params_of(*B), static_cast<double *>(C->d), params_of(*C), C->ncols,
            alpha, beta);
  else
    CLAUNCH(k_cma_scaledadd<float>, C->ncols, static_cast<const float *>(A->d), params_of(*A),
            static_cast<const float *>(B->d), params_of(*B), static_cast<float *>(C->d), params_of(*C), C->ncols,
            static_cast<float>(alpha), static_cast<float>(beta));
  return finish("columnwise sum");
}

int lfric_tri_solve(const lfric_vec *x, const lfric_vec *t0, const lfric_vec *tp, const lfric_vec *tm, lfric_vec *y,
                    lfric_vec *work) {
  if (int rc = require_init()) return rc;
  LF_CHECK(x && t0 && tp && tm && y && work, LFRIC_E_ARG, "null argument");
  const FunctionSpace &f = *x->fs;
  for (const lfric_vec *v : std::initializer_list<const lfric_vec *>{t0, tp, tm, y, work})
    LF_CHECK(v->fs == x->fs && v->precision == x->precision && v->nrhs == 1, LFRIC_E_ARG,
             "tri_solve: fields on different spaces or precisions");
  LF_CHECK(x->nrhs == 1 && f.ndf == 1 && f.top_mask == 0u, LFRIC_E_UNSUPPORTED, "tri_solve: W3 fields expected");
  LF_CHECK(y->d != x->d && work->d != x->d && work->d != y->d, LFRIC_E_ARG, "tri_solve: y and work must be distinct");
  if (x->precision == LFRIC_R64)
    CLAUNCH(k_tri_solve<double>, f.ncells, static_cast<const double *>(x->d), static_cast<const double *>(t0->d),
            static_cast<const double *>(tp->d), static_cast<const double *>(tm->d), static_cast<double *>(y->d),
            static_cast<double *>(work->d), f.d_map, f.ncells, f.nlayers);
  else
    CLAUNCH(k_tri_solve<float>, f.ncells, static_cast<const float *>(x->d), static_cast<const float *>(t0->d),
            static_cast<const float *>(tp->d), static_cast<const float *>(tm->d), static_cast<float *>(y->d),
            static_cast<float *>(work->d), f.d_map, f.ncells, f.nlayers);
  return finish("tri_solve");
}

int lfric_prec_create_cma_inverse(const lfric_cma *A, lfric_prec **out) {
  LF_CHECK(A && out, LFRIC_E_ARG, "null argument");
  LF_CHECK(A->bw == 3 && A->nrow == A->ncol && A->to == A->from, LFRIC_E_UNSUPPORTED,
           "columnwise inverse: a square tridiagonal operator is needed");
  auto *p = new lfric_prec;
  p->kind = 3;
  p->cma = A;
  p->undf = A->to->undf;
  *out = p;
  return LFRIC_OK;
}

}  // extern "C"
