// Abstract linear operators, the vertical-column block-Jacobi preconditioner and the
// device-resident GMRES of liblfric_hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "internal.h"
#include "kernels.h"

namespace lfric {

constexpr int kMaxKrylov = 64;

// Device state of one GMRES restart cycle: Hessenberg matrix (column-major, leading dimension
// m+1), Givens right-hand side and the least-squares solution, all FP64.
struct GmresState {
  double h[(kMaxKrylov + 1) * kMaxKrylov];
  double g[kMaxKrylov + 1];
  double u[kMaxKrylov];
  double cs[kMaxKrylov], sn[kMaxKrylov];  // the cycle's Givens rotations, formed column by column
  double beta;
  double rnorm2;
  double beta0;  // beta of the first cycle
  double ratio;  // |g(iv+1)| / beta0 after column iv: the least-squares residual estimate
};

namespace {

// ---- storage of the block-Jacobi factors ---------------------------------------------------
__device__ inline float bf16_to_f(uint16_t b) { return __uint_as_float(static_cast<uint32_t>(b) << 16); }
__device__ inline uint16_t f_to_bf16(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

template <int STORE>
struct Store;
template <>
struct Store<LFRIC_R32> {
  using T = float;
  __device__ static float val(T v) { return v; }
  __device__ static float get(const T *p, size_t i) { return p[i]; }
  __device__ static void put(T *p, size_t i, float v) { p[i] = v; }
};
template <>
struct Store<LFRIC_BF16> {
  using T = uint16_t;
  __device__ static float val(T v) { return bf16_to_f(v); }
  __device__ static float get(const T *p, size_t i) { return bf16_to_f(p[i]); }
  __device__ static void put(T *p, size_t i, float v) { p[i] = f_to_bf16(v); }
};
template <>
struct Store<LFRIC_F16> {
  using T = _Float16;
  __device__ static float val(T v) { return static_cast<float>(v); }
  __device__ static float get(const T *p, size_t i) { return static_cast<float>(p[i]); }
  __device__ static void put(T *p, size_t i, float v) { p[i] = static_cast<_Float16>(v); }
};

// ---- column tridiagonal of a square LMA operator (colour-ordered, like mm_diagonal) --------
// diag[d] = assembled A(d,d); sup[d] = A(d, d+1) for d and d+1 in one column (0 at the top).
template <typename T, int N>
__global__ __launch_bounds__(kBlock) void k_coltri_lma(const T *__restrict__ A, size_t gs, const int *__restrict__ ebase,
                                                       double *__restrict__ diag, double *__restrict__ sup,
                                                       const int *__restrict__ map, const uint32_t *__restrict__ first,
                                                       int pos0, int npos, int nl, Roles roles) {
  const int H = nl + 1;
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= npos * H) return;
  const int lp = t / H, m = t - lp * H, pos = pos0 + lp;
  const int *mp = map + static_cast<size_t>(pos) * N;
  const uint32_t ft = first[pos];
  const long long e = static_cast<long long>(ebase[pos]) + m;
  for (int i = 0; i < N; ++i) {
    if ((roles.top_mask >> i) & 1u) continue;
    const int pt = roles.partner[i];
    if (m == nl && pt < 0) continue;
    double dg = 0.0, up = 0.0;
    if (m < nl) dg += to_d(A[aidx<T>(gs, e, i * N + i)]);
    if (pt >= 0 && m > 0) dg += to_d(A[aidx<T>(gs, e, pt * N + pt)]);
    if (pt >= 0 && m < nl) up = to_d(A[aidx<T>(gs, e, i * N + pt)]);  // bottom(m) - top(m+1) of layer m
    const size_t o = static_cast<size_t>(mp[i]) + m;
    if ((ft >> i) & 1u) {
      diag[o] = dg;
      sup[o] = up;
    } else {
      diag[o] += dg;
      sup[o] += up;
    }
  }
}

// ---- column tridiagonal of H = M3 + tau D diag(M2)^-1 D^T on W3 (one thread per W3 dof) ----
// W2 dofs of a cell in face order W,S,E,N,B,T; T of layer k is B of layer k+1.
template <typename T>
__global__ void k_coltri_helmholtz(const T *__restrict__ M3, size_t gs3, const T *__restrict__ D, size_t gsd,
                                   const int *__restrict__ ebase, const int *__restrict__ map3,
                                   const int *__restrict__ map2, const T *__restrict__ m2, double tau, int ncells,
                                   int nl, double *__restrict__ diag, double *__restrict__ sup) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<long long>(ncells) * nl) return;
  const int pos = static_cast<int>(t / nl), k = static_cast<int>(t - static_cast<long long>(pos) * nl);
  const long long e = static_cast<long long>(ebase[pos]) + k;
  const int *f = map2 + static_cast<size_t>(pos) * 6;
  double dg = to_d(M3[aidx<T>(gs3, e, 0)]);
  for (int q = 0; q < 6; ++q) {
    const double dq = to_d(D[aidx<T>(gsd, e, q)]);
    dg += tau * dq * dq / to_d(m2[f[q] + k]);
  }
  double up = 0.0;
  if (k < nl - 1) {
    const double dT = to_d(D[aidx<T>(gsd, e, 5)]);
    const double dB = to_d(D[aidx<T>(gsd, e + 1, 4)]);
    up = tau * dT * dB / to_d(m2[f[5] + k]);
  }
  const size_t o = static_cast<size_t>(map3[pos]) + k;
  diag[o] = dg;
  sup[o] = up;
}

// ---- FP32 Thomas factorisation per column, stored in STORE precision -------------------------
template <int STORE>
__global__ void k_coltri_factor(const double *__restrict__ diag, const double *__restrict__ sup,
                                const int *__restrict__ cols, int ncols, typename Store<STORE>::T *__restrict__ fac,
                                long long undf) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const int b = cols[2 * c], h = cols[2 * c + 1];
  using S = Store<STORE>;
  auto *lo = fac, *ipiv = fac + undf, *su = fac + 2 * undf;
  float w_prev = 0.f, s_prev = 0.f;
  for (int k = 0; k < h; ++k) {
    const float d = static_cast<float>(diag[b + k]);
    const float s = k < h - 1 ? static_cast<float>(sup[b + k]) : 0.f;
    const float l = k > 0 ? s_prev / w_prev : 0.f;
    const float w = d - l * s_prev;
    S::put(lo, b + k, l);
    S::put(ipiv, b + k, 1.f / w);
    S::put(su, b + k, s);
    w_prev = w;
    s_prev = s;
  }
}

// ---- apply: z = T^-1 r column by column, FP32 arithmetic ---------------------------------------
template <int STORE, typename V>
__global__ void k_coltri_apply(const typename Store<STORE>::T *__restrict__ fac, long long undf,
                               const int *__restrict__ cols, int ncols, const V *__restrict__ r, V *__restrict__ z) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  const int b = cols[2 * c], h = cols[2 * c + 1];
  using S = Store<STORE>;
  const auto *lo = fac, *ipiv = fac + undf, *su = fac + 2 * undf;
  float y = 0.f;
  for (int k = 0; k < h; ++k) {  // forward: y_k = r_k - l_k y_{k-1}
    y = static_cast<float>(r[b + k]) - S::get(lo, b + k) * y;
    z[b + k] = static_cast<V>(y);
  }
  float zn = 0.f;
  for (int k = h - 1; k >= 0; --k) {  // backward: z_k = (y_k - s_k z_{k+1}) / w_k
    zn = (static_cast<float>(z[b + k]) - S::get(su, b + k) * zn) * S::get(ipiv, b + k);
    z[b + k] = static_cast<V>(zn);
  }
}

// The same solves with the columns sorted by base dof and tiling the dofs, so that the `tile`
// columns of one block are one contiguous dof range: the block's kColThreads threads read the
// range's r and factors into LDS with coalesced loads (kColAhead per thread requested before
// any is stored), the first `tile` lanes solve one column each there (the operations of
// k_coltri_apply, in its order), and the block writes z back coalesced.  cap bounds every
// tile's range.  With vdot the block also leaves the partial sum of vdot . z over its range
// (FP64, as the Gram-Schmidt kernels form theirs) in partials[block]: GMRES's first
// Gram-Schmidt step of the column.
constexpr int kColThreads = 256, kColAhead = 4;

template <int STORE, typename V>
__global__ __launch_bounds__(kColThreads) void k_coltri_apply_tile(const typename Store<STORE>::T *__restrict__ fac,
                                                                   long long undf, const int *__restrict__ cols,
                                                                   int ncols, int tile, int cap,
                                                                   const V *__restrict__ r, V *__restrict__ z,
                                                                   const V *__restrict__ vdot,
                                                                   double *__restrict__ partials) {
  using S = Store<STORE>;
  using FT = typename S::T;
  extern __shared__ __align__(16) unsigned char coltri_lds[];
  float *sv = reinterpret_cast<float *>(coltri_lds);
  FT *sl = reinterpret_cast<FT *>(sv + cap), *si = sl + cap, *ss = si + cap;
  const int c0 = blockIdx.x * tile, c1 = min(c0 + tile, ncols);
  const int lo = cols[2 * c0], len = cols[2 * (c1 - 1)] + cols[2 * (c1 - 1) + 1] - lo;
  for (int i0 = 0; i0 < len; i0 += kColThreads * kColAhead) {
    float rv[kColAhead];
    FT lv[kColAhead], iv[kColAhead], sv4[kColAhead];
#pragma unroll
    for (int a = 0; a < kColAhead; ++a) {
      const int i = i0 + a * kColThreads + static_cast<int>(threadIdx.x);
      const size_t d = static_cast<size_t>(lo) + (i < len ? i : 0);
      rv[a] = static_cast<float>(r[d]);
      lv[a] = fac[d];
      iv[a] = fac[undf + d];
      sv4[a] = fac[2 * undf + d];
    }
#pragma unroll
    for (int a = 0; a < kColAhead; ++a) {
      const int i = i0 + a * kColThreads + static_cast<int>(threadIdx.x);
      if (i < len) {
        sv[i] = rv[a];
        sl[i] = lv[a];
        si[i] = iv[a];
        ss[i] = sv4[a];
      }
    }
  }
  __syncthreads();
  const int c = c0 + static_cast<int>(threadIdx.x);
  if (c < c1) {
    const int b = cols[2 * c] - lo, h = cols[2 * c + 1];
    float y = 0.f;
#pragma unroll 8
    for (int k = 0; k < h; ++k) {
      y = sv[b + k] - S::val(sl[b + k]) * y;
      sv[b + k] = y;
    }
    float zn = 0.f;
#pragma unroll 8
    for (int k = h - 1; k >= 0; --k) {
      zn = (sv[b + k] - S::val(ss[b + k]) * zn) * S::val(si[b + k]);
      sv[b + k] = zn;
    }
  }
  __syncthreads();
  double dot[1] = {0.0};
  for (int i = threadIdx.x; i < len; i += kColThreads) {
    const V zi = static_cast<V>(sv[i]);
    z[static_cast<size_t>(lo) + i] = zi;
    if (vdot) dot[0] += to_d(vdot[static_cast<size_t>(lo) + i]) * to_d(zi);
  }
  if (vdot) block_partials<1>(dot, partials);  // the same for every thread of the block
}

// ---- GMRES -------------------------------------------------------------------------------
// one fused modified Gram-Schmidt step j of Arnoldi column iv:
//   j > 0      : w = w - h(j-1,iv) v_{j-1}
//   j <= iv    : partial v_j . w          -> h(j,iv)
//   j == iv+1  : partial w . w            -> h(iv+1,iv) = sqrt
template <typename T>
__global__ __launch_bounds__(kBlock) void k_mgs_step(T *__restrict__ w, const T *__restrict__ vprev,
                                                     const T *__restrict__ vj, long long n, const double *hprev,
                                                     double *partials) {
  const T a = hprev ? static_cast<T>(-*hprev) : T(0);
  double v[1] = {0.0};
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    T wi = w[i];
    if (vprev) {
      wi = wi + a * vprev[i];
      w[i] = wi;
    }
    v[0] += vj ? to_d(vj[i]) * to_d(wi) : to_d(wi) * to_d(wi);
  }
  block_partials<1>(v, partials);
}

// the sum of a reduction's partials by one block in a fixed order (strided per-thread sums,
// then block_sum's tree); every thread gets it
__device__ inline double finish_sum(const double *__restrict__ partials, int nblocks) {
  double s[1] = {0.0};
  for (int b = threadIdx.x; b < nblocks; b += kBlock) s[0] += partials[b];
  double tot[1];
  block_sum<1>(s, tot);
  return tot[0];
}

__global__ __launch_bounds__(kBlock) void k_finish_to(const double *__restrict__ partials, int nblocks, double *dst,
                                                      int take_sqrt) {
  const double tot = finish_sum(partials, nblocks);
  if (threadIdx.x == 0) *dst = take_sqrt ? sqrt(tot) : tot;
}

// k_mgs_step for j > 0 with the previous step's finish folded in: every block adds the previous
// step's partials itself (finish_sum, the same sum as k_finish_to's) to get h(j-1, iv), block 0
// stores it, then w = w - h(j-1) v_{j-1} and the partials of v_j . w (or w . w) go to partials.
// The thread's first kMgsAhead elements of w, v_{j-1} and v_j are requested before the finish,
// so their latency overlaps it; the elements are taken in the grid-stride order all the same.
constexpr int kMgsAhead = 4;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_mgs_fold(T *__restrict__ w, const T *__restrict__ vprev,
                                                     const T *__restrict__ vj, long long n,
                                                     const double *__restrict__ pprev, int nprev, double *hprev,
                                                     double *partials) {
  const long long i0 = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  T wr[kMgsAhead], pr[kMgsAhead], jr[kMgsAhead];
#pragma unroll
  for (int r = 0; r < kMgsAhead; ++r) {
    const long long i = i0 + r * stride;
    wr[r] = i < n ? w[i] : T(0);
    pr[r] = i < n ? vprev[i] : T(0);
    jr[r] = (vj && i < n) ? vj[i] : T(0);
  }
  const double h = finish_sum(pprev, nprev);
  if (blockIdx.x == 0 && threadIdx.x == 0) *hprev = h;
  const T a = static_cast<T>(-h);
  double v[1] = {0.0};
#pragma unroll
  for (int r = 0; r < kMgsAhead; ++r) {
    const long long i = i0 + r * stride;
    if (i < n) {
      const T wi = wr[r] + a * pr[r];
      w[i] = wi;
      v[0] += vj ? to_d(jr[r]) * to_d(wi) : to_d(wi) * to_d(wi);
    }
  }
  for (long long i = i0 + kMgsAhead * stride; i < n; i += stride) {
    const T wi = w[i] + a * vprev[i];
    w[i] = wi;
    v[0] += vj ? to_d(vj[i]) * to_d(wi) : to_d(wi) * to_d(wi);
  }
  block_partials<1>(v, partials);
}

// The Givens rotations of sci_iterative_solver_mod.f90:929-953 (k_givens below) formed one
// Hessenberg column at a time, as each Arnoldi column is made:
// column iv gets rotations 0 .. iv-1 in order, then its own rotation iv is formed from its
// entries (iv, iv) and (iv+1, iv), applied to it and to g.  Every entry goes through the
// operations of k_givens's loop in the same order (rotation i reaches column j > i before
// rotation i+1 does), so a full cycle ends with the same H and g; a cycle can also stop after
// any column.  Column 0 starts g = (beta, 0, ...); first marks the first cycle, whose beta is
// kept as beta0 for ratio = |g(iv+1)| / beta0 (the preconditioned residual estimate).  Block 0
// of the kernel that finishes hnew = h(iv+1, iv) runs it (k_scale_fold, k_finish_givens): its
// threads read the earlier rotations and the column into LDS together, and thread 0 does the
// arithmetic there and writes the results back.
__device__ inline void givens_column_block(GmresState *st, int iv, int m, int first, double hnew) {
  __shared__ double s_cs[kMaxKrylov], s_sn[kMaxKrylov], s_h[kMaxKrylov + 1];
  double *h = st->h + static_cast<size_t>(iv) * (m + 1), *g = st->g;
  for (int i = threadIdx.x; i < iv; i += blockDim.x) {
    s_cs[i] = st->cs[i];
    s_sn[i] = st->sn[i];
  }
  for (int i = threadIdx.x; i <= iv; i += blockDim.x) s_h[i] = h[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    s_h[iv + 1] = hnew;
    if (iv == 0) {
      g[0] = st->beta;
      for (int i = 1; i <= m; ++i) g[i] = 0.0;
      if (first) st->beta0 = st->beta;
    }
    for (int i = 0; i < iv; ++i) {
      const double ci = s_cs[i], si = s_sn[i];
      const double h1 = ci * s_h[i] + si * s_h[i + 1];
      const double h2 = -si * s_h[i] + ci * s_h[i + 1];
      s_h[i] = h1;
      s_h[i + 1] = h2;
    }
    const double nrm = sqrt(s_h[iv] * s_h[iv] + s_h[iv + 1] * s_h[iv + 1]);
    const double si = s_h[iv + 1] / nrm, ci = s_h[iv] / nrm;
    st->cs[iv] = ci;
    st->sn[iv] = si;
    const double gi = g[iv], gn = g[iv + 1];
    const double p = ci * gi + si * gn;
    const double q = -si * gi + ci * gn;
    g[iv] = p;
    g[iv + 1] = q;
    const double h1 = ci * s_h[iv] + si * s_h[iv + 1];
    const double h2 = -si * s_h[iv] + ci * s_h[iv + 1];
    s_h[iv] = h1;
    s_h[iv + 1] = h2;
    for (int i = 0; i <= iv + 1; ++i) h[i] = s_h[i];
    st->ratio = fabs(q) / st->beta0;
  }
}

// k_finish_to (sqrt) for the cycle's last column, then its Givens column
__global__ __launch_bounds__(kBlock) void k_finish_givens(const double *__restrict__ partials, int nblocks, double *dst,
                                                          GmresState *st, int iv, int m, int first) {
  const double tot = finish_sum(partials, nblocks);
  if (threadIdx.x == 0) *dst = sqrt(tot);
  givens_column_block(st, iv, m, first, sqrt(tot));
}

// k_scale_by_inv with the finish of ||w||^2 folded in: h = sqrt(sum of the partials), block 0
// stores it (and, with st, forms Givens column iv), v = w * (1 / h)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale_fold(T *__restrict__ v, const T *__restrict__ w, long long n,
                                                       const double *__restrict__ partials, int np, double *hout,
                                                       GmresState *st, int iv, int m, int first) {
  const long long i0 = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
  T wr[kMgsAhead];
#pragma unroll
  for (int r = 0; r < kMgsAhead; ++r) wr[r] = i0 + r * stride < n ? w[i0 + r * stride] : T(0);
  const double h = sqrt(finish_sum(partials, np));
  if (blockIdx.x == 0 && threadIdx.x == 0) *hout = h;
  if (st && blockIdx.x == 0) givens_column_block(st, iv, m, first, h);
  const T s = static_cast<T>(1.0 / h);
#pragma unroll
  for (int r = 0; r < kMgsAhead; ++r)
    if (i0 + r * stride < n) v[i0 + r * stride] = wr[r] * s;
  for (long long i = i0 + kMgsAhead * stride; i < n; i += stride) v[i] = w[i] * s;
}

template <typename T>
__global__ void k_div_inplace(T *__restrict__ v, const T *__restrict__ d, long long n) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    v[i] = v[i] / d[i];
}

// v = w * (1 / *h)  (copy then scale, sci_iterative_solver_mod.f90:922-924)
template <typename T>
__global__ void k_scale_by_inv(T *__restrict__ v, const T *__restrict__ w, long long n, const double *h) {
  const T s = static_cast<T>(1.0 / *h);
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    v[i] = w[i] * s;
}

// Givens rotations and back substitution (sci_iterative_solver_mod.f90:929-953), FP64, on the
// device; g(0) = beta, the rest zero.
// The block's threads read H (and g) into LDS together; thread 0 runs the sequential
// arithmetic there and the block writes the results back.
__global__ __launch_bounds__(64) void k_givens(GmresState *st, int m) {
  __shared__ double s_h[(kMaxKrylov + 1) * kMaxKrylov], s_g[kMaxKrylov + 1], s_u[kMaxKrylov];
  const int ld = m + 1;
  for (int i = threadIdx.x; i < ld * m; i += blockDim.x) s_h[i] = st->h[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    double *h = s_h, *g = s_g, *u = s_u;
    g[0] = st->beta;
    for (int i = 1; i <= m; ++i) g[i] = 0.0;
#define H(i, j) h[(j) * ld + (i)]
    for (int iv = 0; iv < m; ++iv) {
      const double nrm = sqrt(H(iv, iv) * H(iv, iv) + H(iv + 1, iv) * H(iv + 1, iv));
      const double si = H(iv + 1, iv) / nrm, ci = H(iv, iv) / nrm;
      const double p = ci * g[iv] + si * g[iv + 1];
      const double q = -si * g[iv] + ci * g[iv + 1];
      g[iv] = p;
      g[iv + 1] = q;
      for (int j = iv; j < m; ++j) {
        const double h1 = ci * H(iv, j) + si * H(iv + 1, j);
        const double h2 = -si * H(iv, j) + ci * H(iv + 1, j);
        H(iv, j) = h1;
        H(iv + 1, j) = h2;
      }
    }
    u[m - 1] = g[m - 1] / H(m - 1, m - 1);
    for (int iv = m - 2; iv >= 0; --iv) {
      double s = g[iv];
      for (int j = iv + 1; j < m; ++j) s = s - H(iv, j) * u[j];
      u[iv] = s / H(iv, iv);
    }
#undef H
  }
  __syncthreads();
  for (int i = threadIdx.x; i < ld * m; i += blockDim.x) st->h[i] = s_h[i];
  for (int i = threadIdx.x; i <= m; i += blockDim.x) st->g[i] = s_g[i];
  for (int i = threadIdx.x; i < m; i += blockDim.x) st->u[i] = s_u[i];
}

// back substitution over the first k columns (k = m after a full cycle), from LDS as k_givens
__global__ __launch_bounds__(64) void k_givens_solve(GmresState *st, int k, int m) {
  __shared__ double s_h[(kMaxKrylov + 1) * kMaxKrylov], s_g[kMaxKrylov], s_u[kMaxKrylov];
  const int ld = m + 1;
  for (int i = threadIdx.x; i < ld * k; i += blockDim.x) s_h[i] = st->h[i];
  for (int i = threadIdx.x; i < k; i += blockDim.x) s_g[i] = st->g[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    const double *h = s_h, *g = s_g;
    double *u = s_u;
#define H(i, j) h[(j) * ld + (i)]
    u[k - 1] = g[k - 1] / H(k - 1, k - 1);
    for (int iv = k - 2; iv >= 0; --iv) {
      double s = g[iv];
      for (int j = iv + 1; j < k; ++j) s = s - H(iv, j) * u[j];
      u[iv] = s / H(iv, iv);
    }
#undef H
  }
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += blockDim.x) st->u[i] = s_u[i];
}

// x = x + sum_iv u(iv) v_iv, iv in order (sci_iterative_solver_mod.f90:956-960)
template <typename T>
__global__ void k_gmres_update(T *__restrict__ x, const T *__restrict__ V, long long n, int m, const double *u) {
  for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    T xi = x[i];
    for (int iv = 0; iv < m; ++iv) xi = xi + static_cast<T>(u[iv]) * V[static_cast<size_t>(iv) * n + i];
    x[i] = xi;
  }
}

}  // namespace

// ======================================================================================
static size_t esz(int p) { return p == LFRIC_R64 ? 8 : p == LFRIC_R32 ? 4 : 2; }

#define LAUNCHK(kernel, grid, block, ...) hipLaunchKernelGGL((kernel), dim3(grid), dim3(block), 0, cur_stream(), __VA_ARGS__)

int linop_apply(const lfric_linop *L, const void *x, void *y) {
  if (L->kind == 0) {
    if (int rc = launch_lma_apply_solver(L->A, x, y, nullptr)) return rc;
    return L->apply_bc ? launch_enforce_bc(*L->space, L->precision, y, nullptr) : LFRIC_OK;
  }
  // y = M3 x + tau * D (D^T x / m2)
  if (L->fused < 0) L->fused = helmholtz_fusable(L) ? 1 : 0;
  if (L->fused) return launch_helmholtz_fused(L, x, y);
  const long long n2 = L->D->from->undf, n3 = L->space->undf;
  if (int rc = launch_lma_transpose(L->D, x, L->t_w2)) return rc;
  if (L->precision == LFRIC_R64)
    LAUNCHK(k_div_inplace<double>, grid_for(n2, kBlock, 8192), kBlock, static_cast<double *>(L->t_w2),
            static_cast<const double *>(L->m2->d), n2);
  else
    LAUNCHK(k_div_inplace<float>, grid_for(n2, kBlock, 8192), kBlock, static_cast<float *>(L->t_w2),
            static_cast<const float *>(L->m2->d), n2);
  LF_HIP(hipGetLastError());
  if (int rc = launch_dg_apply(L->D, L->t_w2, L->t_w3)) return rc;
  if (int rc = launch_dg_apply(L->M3, x, y)) return rc;
  return launch_axpy(L->precision, y, L->t_w3, n3, L->tau);
}

// the tiled column block-Jacobi's blocks (one partial sum each when it also forms a dot), 0
// when P does not apply that way
static int coltri_tiles(const lfric_prec *P) {
  return (P && P->kind == 2 && P->tile) ? (P->ncols + P->tile - 1) / P->tile : 0;
}

// y = P^-1 x; vdot/partials only with coltri_tiles(P) > 0: also the partials of vdot . y
static int prec_apply_impl(const lfric_prec *P, int prec, const void *x, void *y, long long n, const void *vdot,
                           double *partials) {
  if (!P || P->kind == 0) return launch_copy(prec, y, x, n);
  if (P->kind == 1) {
    LF_CHECK(P->diag->precision == prec && P->diag->n == n, LFRIC_E_ARG, "diagonal does not match vector");
    return launch_divide(prec, y, x, P->diag->d, n);
  }
  LF_CHECK(P->undf == n, LFRIC_E_ARG, "column preconditioner does not match vector");
  if (P->kind == 3) {
    LF_CHECK(P->cma->precision == prec, LFRIC_E_ARG, "columnwise preconditioner precision differs from the vector");
    return launch_cma_appinv(P->cma, x, y);
  }
  const int grid = grid_for(P->ncols, 64);
  const int tgrid = P->tile ? (P->ncols + P->tile - 1) / P->tile : 0;
#define CT_APPLY(ST)                                                                                           \
  if (P->storage == ST) {                                                                                      \
    using FT = typename Store<ST>::T;                                                                          \
    const size_t lds = static_cast<size_t>(P->cap) * (sizeof(float) + 3 * sizeof(FT));                       \
    if (P->tile && prec == LFRIC_R64)                                                                          \
      hipLaunchKernelGGL((k_coltri_apply_tile<ST, double>), dim3(tgrid), dim3(kColThreads), lds,               \
                         cur_stream(), static_cast<const FT *>(P->d_fac), P->undf, P->d_cols, P->ncols,        \
                         P->tile, P->cap, static_cast<const double *>(x), static_cast<double *>(y),            \
                         static_cast<const double *>(vdot), partials);                                         \
    else if (P->tile)                                                                                          \
      hipLaunchKernelGGL((k_coltri_apply_tile<ST, float>), dim3(tgrid), dim3(kColThreads), lds,                \
                         cur_stream(), static_cast<const FT *>(P->d_fac), P->undf, P->d_cols, P->ncols,        \
                         P->tile, P->cap, static_cast<const float *>(x), static_cast<float *>(y),              \
                         static_cast<const float *>(vdot), partials);                                          \
    else if (prec == LFRIC_R64)                                                                                \
      LAUNCHK((k_coltri_apply<ST, double>), grid, 64, static_cast<const FT *>(P->d_fac), P->undf, P->d_cols,   \
              P->ncols, static_cast<const double *>(x), static_cast<double *>(y));                             \
    else                                                                                                       \
      LAUNCHK((k_coltri_apply<ST, float>), grid, 64, static_cast<const FT *>(P->d_fac), P->undf, P->d_cols,    \
              P->ncols, static_cast<const float *>(x), static_cast<float *>(y));                               \
    LF_HIP(hipGetLastError());                                                                                 \
    return LFRIC_OK;                                                                                           \
  }
  CT_APPLY(LFRIC_R32)
  CT_APPLY(LFRIC_BF16)
  CT_APPLY(LFRIC_F16)
#undef CT_APPLY
  return fail(LFRIC_E_UNSUPPORTED, "column preconditioner storage");
}

int prec_apply_any(const lfric_prec *P, int prec, const void *x, void *y, long long n) {
  return prec_apply_impl(P, prec, x, y, n, nullptr, nullptr);
}

// scratch of the GMRES driver, grown on demand
static void *g_gm[6] = {nullptr};
static size_t g_gm_bytes[6] = {0};
static GmresState *g_gstate = nullptr;
static double *g_gpart = nullptr;
static double *g_host_scalar = nullptr;
static hipEvent_t g_gev[2] = {nullptr, nullptr};  // the inner solve's per-column residual reads
constexpr int kGmBlocks = 1024;

static int gm_ws(int slot, size_t bytes, void **out) {
  if (g_gm_bytes[slot] < bytes) {
    dfree(g_gm[slot], g_gm_bytes[slot]);
    g_gm[slot] = nullptr;
    g_gm_bytes[slot] = 0;
    if (int rc = dmalloc(&g_gm[slot], bytes)) return rc;
    g_gm_bytes[slot] = bytes;
  }
  *out = g_gm[slot];
  return LFRIC_OK;
}

static int gm_init() {
  // each piece on its own, so that a call after a failed one finishes the set
  if (!g_gstate)
    if (int rc = dmalloc(reinterpret_cast<void **>(&g_gstate), sizeof(GmresState))) return rc;
  if (!g_gpart)
    if (int rc = dmalloc(reinterpret_cast<void **>(&g_gpart), sizeof(double) * 2 * kGmBlocks)) return rc;
  if (!g_host_scalar)
    LF_HIP(hipHostMalloc(reinterpret_cast<void **>(&g_host_scalar), sizeof(double) * 4, hipHostMallocDefault));
  for (auto &e : g_gev)
    if (!e) LF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return LFRIC_OK;
}

// The Gram-Schmidt steps j = 0 .. iv+1 of Arnoldi column iv (w against v_0 .. v_iv, then
// ||w||), and v_{iv+1} = w / h(iv+1, iv) unless iv is the cycle's last column.  Each step's
// partial sums are finished by the next kernel's blocks themselves (partials in two alternating
// halves of the scratch), so a step is one launch; the h values are those of separate
// finishing kernels bit for bit.
// first_parts > 0: step 0's partials (v_0 . w) are already in the first half of the scratch,
// that many of them (left there by the preconditioner's apply).
template <typename T>
static int arnoldi_mgs(T *w, T *V, long long n, int iv, int m, double *h, GmresState *gst = nullptr,
                       int first = 0, int first_parts = 0) {
  const int vg = grid_for(n);
  double *part[2] = {g_gpart, g_gpart + kGmBlocks};
  if (!first_parts)
    LAUNCHK(k_mgs_step<T>, kGmBlocks, kBlock, w, static_cast<const T *>(nullptr), static_cast<const T *>(V), n,
            static_cast<const double *>(nullptr), part[0]);
  for (int j = 1; j <= iv + 1; ++j) {
    const T *vj = j <= iv ? V + static_cast<size_t>(j) * n : nullptr;
    const int nprev = (j == 1 && first_parts) ? first_parts : kGmBlocks;
    LAUNCHK(k_mgs_fold<T>, kGmBlocks, kBlock, w, V + static_cast<size_t>(j - 1) * n, vj, n, part[(j - 1) & 1],
            nprev, &h[iv * (m + 1) + (j - 1)], part[j & 1]);
  }
  double *hlast = &h[iv * (m + 1) + iv + 1];
  if (iv < m - 1)
    LAUNCHK(k_scale_fold<T>, std::min(vg, kGmBlocks), kBlock, V + static_cast<size_t>(iv + 1) * n, w, n,
            part[(iv + 1) & 1], kGmBlocks, hlast, gst, iv, m, first);
  else if (gst)
    LAUNCHK(k_finish_givens, 1, kBlock, part[(iv + 1) & 1], kGmBlocks, hlast, gst, iv, m, first);
  else
    LAUNCHK(k_finish_to, 1, kBlock, part[(iv + 1) & 1], kGmBlocks, hlast, 1);
  LF_HIP(hipGetLastError());
  return LFRIC_OK;
}

// ||v||^2 or v.w into *dst (device) with the fused MGS kernel
template <typename T>
static int dot_into(const T *a, T *b, long long n, double *dst, bool sq) {
  LAUNCHK(k_mgs_step<T>, kGmBlocks, kBlock, b, static_cast<const T *>(nullptr), sq ? static_cast<const T *>(nullptr) : a,
          n, static_cast<const double *>(nullptr), g_gpart);
  LAUNCHK(k_finish_to, 1, kBlock, g_gpart, kGmBlocks, dst, sq ? 1 : 0);
  LF_HIP(hipGetLastError());
  return LFRIC_OK;
}

static int read_scalar(const double *d, double *h) {
  LF_HIP(hipMemcpyAsync(g_host_scalar, d, sizeof(double), hipMemcpyDeviceToHost, cur_stream()));
  LF_HIP(hipStreamSynchronize(cur_stream()));
  *h = *g_host_scalar;
  return LFRIC_OK;
}

struct GmresOut {
  int cycles = 0, iterations = 0, status = LFRIC_NOT_CONVERGED;
  double r0 = 0, rn = 0;
};

// gmres_solve (sci_iterative_solver_mod.f90:830-1003) with every vector on the device.  One
// host synchronisation per restart cycle (the convergence test on the true residual).
//
// early (the inner solve of the refinement only; the reference's GMRES runs whole cycles):
// after each column the host also reads the least-squares residual estimate |g(iv+1)| / beta0,
// one column behind the device (the read of column iv-1 waits while column iv runs), and ends
// the cycle with the columns made so far once it falls below r_tol, and the solve with it (the
// refinement's FP64 residual is the test that follows); a cycle that runs to m columns ends
// with the true-residual test as usual.
template <typename T>
static int gmres_t(const lfric_linop *A, const lfric_prec *P, T *x, const T *b, long long n, int m, double r_tol,
                   double a_tol, int max_cycles, int monitor, GmresOut &out, bool early = false) {
  if (int rc = gm_init()) return rc;
  LF_CHECK(m >= 1 && m <= kMaxKrylov, LFRIC_E_ARG, "gcrk must be in [1, 64]");
  const int prec = sizeof(T) == 8 ? LFRIC_R64 : LFRIC_R32;
  void *vV, *vw, *vs, *vr, *vAx;
  if (int rc = gm_ws(0, sizeof(T) * static_cast<size_t>(n) * m, &vV)) return rc;
  if (int rc = gm_ws(1, sizeof(T) * static_cast<size_t>(n), &vw)) return rc;
  if (int rc = gm_ws(2, sizeof(T) * static_cast<size_t>(n), &vs)) return rc;
  if (int rc = gm_ws(3, sizeof(T) * static_cast<size_t>(n), &vr)) return rc;
  if (int rc = gm_ws(4, sizeof(T) * static_cast<size_t>(n), &vAx)) return rc;
  T *V = static_cast<T *>(vV), *w = static_cast<T *>(vw), *s = static_cast<T *>(vs), *r = static_cast<T *>(vr);
  T *Ax = static_cast<T *>(vAx);
  GmresState *st = g_gstate;
  const int vg = grid_for(n, kBlock, 8192);
  // res = b - A x; sc_err = ||res||
  if (int rc = linop_apply(A, x, Ax)) return rc;
  if (int rc = launch_copy(prec, r, b, n)) return rc;
  if (int rc = launch_axpy(prec, r, Ax, n, -1.0)) return rc;
  if (int rc = dot_into<T>(r, r, n, &st->rnorm2, true)) return rc;
  double sc_err;
  if (int rc = read_scalar(&st->rnorm2, &sc_err)) return rc;
  out.r0 = sc_err;
  out.rn = sc_err;
  if (!(sc_err == sc_err)) {
    out.status = LFRIC_NAN;
    return LFRIC_OK;
  }
  if (sc_err < 2.220446049250313e-16) {
    out.status = LFRIC_OK;
    return LFRIC_OK;
  }
  if (monitor) sc_err = std::max(sc_err, a_tol);
  for (int cyc = 1; cyc <= max_cycles; ++cyc) {
    if (int rc = prec_apply_any(P, prec, r, s, n)) return rc;
    // beta = ||s|| and v_0 = s / beta: the partials of s . s, then one kernel that finishes them
    // (as k_finish_to) and scales (as k_scale_by_inv)
    LAUNCHK(k_mgs_step<T>, kGmBlocks, kBlock, s, static_cast<const T *>(nullptr), static_cast<const T *>(nullptr), n,
            static_cast<const double *>(nullptr), g_gpart);
    LAUNCHK(k_scale_fold<T>, std::min(grid_for(n), kGmBlocks), kBlock, V, s, n, g_gpart, kGmBlocks, &st->beta,
            static_cast<GmresState *>(nullptr), 0, m, 0);
    int k = m;  // columns made in this cycle
    for (int iv = 0; iv < m; ++iv) {
      T *viv = V + static_cast<size_t>(iv) * n;
      int first_parts = 0;
      if (!P || P->kind == 0) {  // identity: w = P^-1 (A v_iv) is the apply's output itself
        if (int rc = linop_apply(A, viv, w)) return rc;
      } else if (coltri_tiles(P) > 0 && coltri_tiles(P) <= kGmBlocks) {
        // the tiled column block-Jacobi also leaves Gram-Schmidt step 0's partials (v_0 . w)
        if (int rc = linop_apply(A, viv, s)) return rc;
        if (int rc = prec_apply_impl(P, prec, s, w, n, V, g_gpart)) return rc;
        first_parts = coltri_tiles(P);
      } else {
        if (int rc = linop_apply(A, viv, s)) return rc;  // s = A w (w = v_iv)
        if (int rc = prec_apply_any(P, prec, s, w, n)) return rc;  // w = P^-1 s
      }
      // Gram-Schmidt and Givens column iv
      if (int rc = arnoldi_mgs<T>(w, V, n, iv, m, st->h, st, cyc == 1 ? 1 : 0, first_parts)) return rc;
      if (early && iv < m - 1) {
        double *slot = g_host_scalar + 2 + (iv & 1);
        LF_HIP(hipMemcpyAsync(slot, &st->ratio, sizeof(double), hipMemcpyDeviceToHost, cur_stream()));
        LF_HIP(hipEventRecord(g_gev[iv & 1], cur_stream()));
        if (iv >= 1) {
          LF_HIP(hipEventSynchronize(g_gev[(iv - 1) & 1]));
          if (g_host_scalar[2 + ((iv - 1) & 1)] < r_tol) {  // column iv is made too: use it
            k = iv + 1;
            break;
          }
        }
      }
    }
    out.iterations += k;
    LAUNCHK(k_givens_solve, 1, 64, st, k, m);
    LAUNCHK(k_gmres_update<T>, vg, kBlock, x, V, n, k, st->u);
    LF_HIP(hipGetLastError());
    if (k < m) {  // ended on the estimate: the refinement's own residual is the test
      out.cycles = cyc;
      out.status = LFRIC_OK;
      return LFRIC_OK;
    }
    if (int rc = linop_apply(A, x, Ax)) return rc;
    if (int rc = launch_copy(prec, r, Ax, n)) return rc;
    if (int rc = launch_cg_residual(prec, r, b, n)) return rc;  // res%aypx(-1, b)
    if (int rc = dot_into<T>(r, r, n, &st->rnorm2, true)) return rc;
    double rn;
    if (int rc = read_scalar(&st->rnorm2, &rn)) return rc;
    out.cycles = cyc;
    out.rn = rn;
    if (!(rn == rn)) {
      out.status = LFRIC_NAN;
      return LFRIC_OK;
    }
    if (monitor && rn / sc_err < r_tol) {
      out.status = LFRIC_OK;
      return LFRIC_OK;
    }
  }
  out.status = monitor ? LFRIC_NOT_CONVERGED : LFRIC_OK;
  return LFRIC_OK;
}

// fgmres_solve (sci_iterative_solver_mod.f90:1055-1237): flexible (right-preconditioned)
// restarted GMRES.  As in the reference the residual starts from b (a zero initial guess is
// assumed, :1080), the correction dx is built from the preconditioned basis Pv, and x += dx at
// the end.  One host synchronisation per restart cycle.
template <typename T>
static int fgmres_t(const lfric_linop *A, const lfric_prec *P, T *x, const T *b, long long n, int m, double r_tol,
                    double a_tol, int max_cycles, int monitor, GmresOut &out) {
  if (int rc = gm_init()) return rc;
  LF_CHECK(m >= 1 && m <= kMaxKrylov, LFRIC_E_ARG, "gcrk must be in [1, 64]");
  const int prec = sizeof(T) == 8 ? LFRIC_R64 : LFRIC_R32;
  void *vV, *vPV, *vw, *vr, *vdx;
  if (int rc = gm_ws(0, sizeof(T) * static_cast<size_t>(n) * m, &vV)) return rc;
  if (int rc = gm_ws(5, sizeof(T) * static_cast<size_t>(n) * m, &vPV)) return rc;
  if (int rc = gm_ws(1, sizeof(T) * static_cast<size_t>(n), &vw)) return rc;
  if (int rc = gm_ws(3, sizeof(T) * static_cast<size_t>(n), &vr)) return rc;
  if (int rc = gm_ws(4, sizeof(T) * static_cast<size_t>(n), &vdx)) return rc;
  T *V = static_cast<T *>(vV), *PV = static_cast<T *>(vPV), *w = static_cast<T *>(vw);
  T *r = static_cast<T *>(vr), *dx = static_cast<T *>(vdx);
  GmresState *st = g_gstate;
  const int vg = grid_for(n, kBlock, 8192);
  if (int rc = launch_set(prec, dx, n, 0.0)) return rc;
  if (int rc = launch_copy(prec, r, b, n)) return rc;  // res = b
  if (int rc = dot_into<T>(r, r, n, &st->rnorm2, true)) return rc;
  double sc_err;
  if (int rc = read_scalar(&st->rnorm2, &sc_err)) return rc;
  out.r0 = out.rn = sc_err;
  if (!(sc_err == sc_err)) {
    out.status = LFRIC_NAN;
    return LFRIC_OK;
  }
  if (sc_err < 2.220446049250313e-16) {
    out.status = LFRIC_OK;
    return LFRIC_OK;
  }
  if (monitor) sc_err = std::max(sc_err, a_tol);
  // beta = ||s||, v1 = s / beta (s = res)
  if (int rc = dot_into<T>(r, r, n, &st->beta, true)) return rc;
  LAUNCHK(k_scale_by_inv<T>, vg, kBlock, V, r, n, &st->beta);
  out.status = monitor ? LFRIC_NOT_CONVERGED : LFRIC_OK;
  for (int cyc = 1; cyc <= max_cycles; ++cyc) {
    for (int iv = 0; iv < m; ++iv) {
      T *viv = V + static_cast<size_t>(iv) * n, *pviv = PV + static_cast<size_t>(iv) * n;
      if (int rc = prec_apply_any(P, prec, viv, pviv, n)) return rc;  // Pv = P^-1 v
      if (int rc = linop_apply(A, pviv, w)) return rc;               // w = s = A Pv
      if (int rc = arnoldi_mgs<T>(w, V, n, iv, m, st->h)) return rc;
    }
    LAUNCHK(k_givens, 1, 64, st, m);
    LAUNCHK(k_gmres_update<T>, vg, kBlock, dx, PV, n, m, st->u);  // dx += sum u_iv Pv_iv
    LF_HIP(hipGetLastError());
    if (int rc = linop_apply(A, dx, w)) return rc;                 // Ax = A dx
    if (int rc = launch_copy(prec, r, w, n)) return rc;
    if (int rc = launch_cg_residual(prec, r, b, n)) return rc;     // res%aypx(-1, b)
    if (int rc = dot_into<T>(r, r, n, &st->beta, true)) return rc;
    double rn;
    if (int rc = read_scalar(&st->beta, &rn)) return rc;
    out.cycles = cyc;
    out.rn = rn;
    if (!(rn == rn)) {
      out.status = LFRIC_NAN;
      break;
    }
    if (monitor && rn / sc_err < r_tol) {
      out.status = LFRIC_OK;
      break;
    }
    LAUNCHK(k_scale_by_inv<T>, vg, kBlock, V, r, n, &st->beta);  // v1 = res / beta
  }
  return launch_axpy(prec, x, dx, n, 1.0);  // x = x + dx
}

}  // namespace lfric

using namespace lfric;

lfric_prec::~lfric_prec() {
  dfree(d_cols, cols_bytes);
  dfree(d_fac, fac_bytes);
}

lfric_linop::~lfric_linop() {
  dfree(d_hcol, hcol_bytes);
  dfree(d_dtc, dtc_bytes);
  dfree(d_drc, drc_bytes);
  dfree(t_w2, t_w2_bytes);
  dfree(t_w3, t_w3_bytes);
}

extern "C" {

int lfric_linop_create_lma(const lfric_op *A, lfric_linop **out) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A && out, LFRIC_E_ARG, "null argument");
  LF_CHECK(A->to == A->from, LFRIC_E_ARG, "a linear operator for the solvers must be square on one space");
  auto *L = new lfric_linop;
  L->kind = 0;
  L->A = A;
  L->precision = A->precision;
  L->space = A->to;
  *out = L;
  return LFRIC_OK;
}

int lfric_linop_create_mass_matrix(const lfric_op *A, int apply_bc, lfric_linop **out) {
  if (int rc = lfric_linop_create_lma(A, out)) return rc;
  (*out)->apply_bc = apply_bc ? 1 : 0;
  return LFRIC_OK;
}

int lfric_linop_create_helmholtz(const lfric_op *M3, const lfric_op *D, const lfric_vec *m2, double tau,
                                 lfric_linop **out) {
  if (int rc = require_init()) return rc;
  LF_CHECK(M3 && D && m2 && out, LFRIC_E_ARG, "null argument");
  LF_CHECK(M3->to == M3->from && D->to == M3->to && m2->fs == D->from, LFRIC_E_ARG,
           "Helmholtz operator: M3 on W3, D from W2 to W3, m2_diag on W2 expected");
  LF_CHECK(M3->precision == D->precision && m2->precision == D->precision, LFRIC_E_ARG, "mixed precisions");
  LF_CHECK(D->to->ndf == 1 && D->from->ndf == 6, LFRIC_E_UNSUPPORTED, "Helmholtz operator needs W3 (1 dof) and W2 (6 dofs)");
  auto *L = new lfric_linop;
  L->kind = 1;
  L->M3 = M3;
  L->D = D;
  L->m2 = m2;
  L->tau = tau;
  L->precision = D->precision;
  L->space = M3->to;
  L->t_w2_bytes = esz(L->precision) * static_cast<size_t>(D->from->undf);
  L->t_w3_bytes = esz(L->precision) * static_cast<size_t>(M3->to->undf);
  if (int rc = dmalloc(&L->t_w2, L->t_w2_bytes)) {
    delete L;
    return rc;
  }
  if (int rc = dmalloc(&L->t_w3, L->t_w3_bytes)) {
    delete L;
    return rc;
  }
  *out = L;
  return LFRIC_OK;
}

int lfric_linop_apply(const lfric_linop *L, const lfric_vec *x, lfric_vec *y) {
  if (int rc = require_init()) return rc;
  LF_CHECK(L && x && y && x->fs == L->space && y->fs == L->space, LFRIC_E_ARG, "vectors must live on the operator space");
  LF_CHECK(x->precision == L->precision && y->precision == L->precision && x->d != y->d, LFRIC_E_ARG,
           "vector precision differs from the operator or x aliases y");
  return linop_apply(L, x->d, y->d);
}

void lfric_linop_destroy(lfric_linop *L) { delete L; }

int lfric_prec_create_column_jacobi(const lfric_linop *L, int storage, lfric_prec **out) {
  if (int rc = require_init()) return rc;
  LF_CHECK(L && out, LFRIC_E_ARG, "null argument");
  LF_CHECK(storage == LFRIC_R32 || storage == LFRIC_BF16 || storage == LFRIC_F16, LFRIC_E_UNSUPPORTED,
           "column preconditioner storage must be R32, BF16 or F16");
  const FunctionSpace &f = *L->space;
  const long long n = f.undf;
  double *diag = nullptr, *sup = nullptr;
  if (int rc = dmalloc(reinterpret_cast<void **>(&diag), sizeof(double) * n)) return rc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&sup), sizeof(double) * n)) return rc;
  // columns: one per bottom / unpaired dof the first time its column is met
  std::vector<int> cols;
  {
    std::vector<char> seen(static_cast<size_t>(n) + 1, 0);
    for (int p = 0; p < f.ncells; ++p) {
      const int cell = f.cell_of_pos[p];
      for (int d = 0; d < f.ndf; ++d) {
        if ((f.top_mask >> d) & 1u) continue;
        const int base = f.dofmap[static_cast<size_t>(cell) * f.ndf + d] - 1;
        if (seen[base]) continue;
        seen[base] = 1;
        cols.push_back(base);
        cols.push_back(f.nlayers + (f.partner[d] >= 0 ? 1 : 0));
      }
    }
    // by base dof: each column's solve is independent of the others, so the order only decides
    // which columns share a block of the tiled apply
    std::vector<std::pair<int, int>> sorted(cols.size() / 2);
    for (size_t i = 0; i < sorted.size(); ++i) sorted[i] = {cols[2 * i], cols[2 * i + 1]};
    std::sort(sorted.begin(), sorted.end());
    for (size_t i = 0; i < sorted.size(); ++i) {
      cols[2 * i] = sorted[i].first;
      cols[2 * i + 1] = sorted[i].second;
    }
  }
  int rc = LFRIC_OK;
  if (L->kind == 0) {
    const lfric_op *A = L->A;
    const Roles roles = roles_of(f);
    for (int c = 0; c < f.ncolours && rc == LFRIC_OK; ++c) {
      const int pos0 = f.colour_start[c], npos = f.colour_start[c + 1] - pos0;
      if (!npos) continue;
      const int grid = grid_for(static_cast<long long>(npos) * (f.nlayers + 1));
#define CT_CASE(N)                                                                                                  \
  if (f.ndf == N) {                                                                                                 \
    if (A->precision == LFRIC_R64)                                                                                  \
      LAUNCHK((k_coltri_lma<double, N>), grid, kBlock, static_cast<const double *>(A->d), A->gs, f.d_ebase, diag,   \
              sup, f.d_map, f.d_first, pos0, npos, f.nlayers, roles);                                               \
    else                                                                                                            \
      LAUNCHK((k_coltri_lma<float, N>), grid, kBlock, static_cast<const float *>(A->d), A->gs, f.d_ebase, diag,     \
              sup, f.d_map, f.d_first, pos0, npos, f.nlayers, roles);                                               \
    continue;                                                                                                       \
  }
      CT_CASE(8)
      CT_CASE(27)
      CT_CASE(6)
      CT_CASE(1)
#undef CT_CASE
      rc = fail(LFRIC_E_UNSUPPORTED, "column preconditioner: unsupported ndf");
    }
  } else {
    const FunctionSpace &w2 = *L->D->from;
    const long long nthreads = static_cast<long long>(f.ncells) * f.nlayers;
    if (L->precision == LFRIC_R64)
      LAUNCHK(k_coltri_helmholtz<double>, grid_for(nthreads), kBlock, static_cast<const double *>(L->M3->d), L->M3->gs,
              static_cast<const double *>(L->D->d), L->D->gs, f.d_ebase, f.d_map, w2.d_map,
              static_cast<const double *>(L->m2->d), L->tau, f.ncells, f.nlayers, diag, sup);
    else
      LAUNCHK(k_coltri_helmholtz<float>, grid_for(nthreads), kBlock, static_cast<const float *>(L->M3->d), L->M3->gs,
              static_cast<const float *>(L->D->d), L->D->gs, f.d_ebase, f.d_map, w2.d_map,
              static_cast<const float *>(L->m2->d), L->tau, f.ncells, f.nlayers, diag, sup);
  }
  auto *P = new lfric_prec;
  P->kind = 2;
  P->storage = storage;
  P->undf = n;
  P->ncols = static_cast<int>(cols.size() / 2);
  // tiles of up to 32 consecutive columns when the sorted columns tile [0, undf), as many as
  // fit 48 KB of LDS at the tallest column; otherwise one thread per column straight from memory
  {
    constexpr size_t kColLds = 48 * 1024;
    const size_t per_dof = 4 + 3 * esz(storage);
    bool tiles = P->ncols > 0;
    long long next = 0;
    int hmax = 0;
    for (size_t i = 0; i < cols.size(); i += 2) {
      tiles = tiles && cols[i] == next;
      next = static_cast<long long>(cols[i]) + cols[i + 1];
      hmax = std::max(hmax, cols[i + 1]);
    }
    tiles = tiles && next == n && hmax > 0;
    const int tile = tiles ? static_cast<int>(std::min<size_t>(32, kColLds / (per_dof * hmax))) : 0;
    int cap = 0;
    for (int c0 = 0; tile > 0 && c0 < P->ncols; c0 += tile) {
      const int c1 = std::min(c0 + tile, P->ncols);
      cap = std::max(cap, cols[2 * (c1 - 1)] + cols[2 * (c1 - 1) + 1] - cols[2 * c0]);
    }
    if (tile > 0 && static_cast<size_t>(cap) * per_dof <= kColLds) {
      P->tile = tile;
      P->cap = cap;
    }
  }
  P->cols_bytes = sizeof(int) * cols.size();
  P->fac_bytes = 3 * esz(storage) * static_cast<size_t>(n);
  if (rc == LFRIC_OK) rc = dmalloc(reinterpret_cast<void **>(&P->d_cols), P->cols_bytes);
  if (rc == LFRIC_OK) rc = dmalloc(&P->d_fac, P->fac_bytes);
  if (rc == LFRIC_OK) {
    hipError_t e = hipMemcpyAsync(P->d_cols, cols.data(), P->cols_bytes, hipMemcpyHostToDevice, cur_stream());
    if (e != hipSuccess) rc = fail(LFRIC_E_HIP, hipGetErrorString(e));
  }
  if (rc == LFRIC_OK) {
    const int grid = grid_for(P->ncols, 64);
    if (storage == LFRIC_R32)
      LAUNCHK(k_coltri_factor<LFRIC_R32>, grid, 64, diag, sup, P->d_cols, P->ncols, static_cast<float *>(P->d_fac), n);
    else if (storage == LFRIC_BF16)
      LAUNCHK(k_coltri_factor<LFRIC_BF16>, grid, 64, diag, sup, P->d_cols, P->ncols, static_cast<uint16_t *>(P->d_fac), n);
    else
      LAUNCHK(k_coltri_factor<LFRIC_F16>, grid, 64, diag, sup, P->d_cols, P->ncols, static_cast<_Float16 *>(P->d_fac), n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(cur_stream());
    if (e != hipSuccess) rc = fail(LFRIC_E_HIP, hipGetErrorString(e));
  }
  dfree(diag, sizeof(double) * n);
  dfree(sup, sizeof(double) * n);
  if (rc != LFRIC_OK) {
    delete P;
    return rc;
  }
  *out = P;
  return LFRIC_OK;
}

static int gmres_common(const lfric_linop *A, const lfric_prec *P, void *x, const void *b, long long n, int prec,
                        const lfric_solver_opts *o, GmresOut &g, bool early = false) {
  if (prec == LFRIC_R64)
    return gmres_t<double>(A, P, static_cast<double *>(x), static_cast<const double *>(b), n, o->gcrk, o->r_tol,
                           o->a_tol, o->max_iter, o->monitor_convergence, g, early);
  return gmres_t<float>(A, P, static_cast<float *>(x), static_cast<const float *>(b), n, o->gcrk, o->r_tol, o->a_tol,
                        o->max_iter, o->monitor_convergence, g, early);
}

int lfric_gmres_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                      const lfric_solver_opts *opts, lfric_solver_result *res) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A && x && b && opts, LFRIC_E_ARG, "null argument");
  LF_CHECK(x->fs == A->space && b->fs == A->space && x->precision == A->precision && b->precision == A->precision,
           LFRIC_E_ARG, "vectors must match the operator space and precision");
  GmresOut g;
  if (int rc = gmres_common(A, P, x->d, b->d, x->n, A->precision, opts, g)) return rc;
  if (res) {
    res->status = g.status;
    res->iterations = g.cycles * opts->gcrk;
    res->outer_iterations = g.cycles;
    res->r_nrm_0 = g.r0;
    res->r_nrm = g.rn;
    res->true_rel_res = g.r0 > 0 ? g.rn / g.r0 : 0.0;
  }
  if (g.status == LFRIC_NOT_CONVERGED && !opts->fail_on_non_converged) return LFRIC_OK;
  if (g.status != LFRIC_OK) set_error(g.status == LFRIC_NAN ? "NaN in GMRES residual" : "GMRES did not converge");
  return g.status;
}

int lfric_fgmres_solve(const lfric_linop *A, const lfric_prec *P, lfric_vec *x, const lfric_vec *b,
                       const lfric_solver_opts *opts, lfric_solver_result *res) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A && x && b && opts, LFRIC_E_ARG, "null argument");
  LF_CHECK(x->fs == A->space && b->fs == A->space && x->precision == A->precision && b->precision == A->precision,
           LFRIC_E_ARG, "vectors must match the operator space and precision");
  LF_CHECK(x->d != b->d, LFRIC_E_ARG, "x and b must differ");
  GmresOut g;
  int rc = A->precision == LFRIC_R64
               ? fgmres_t<double>(A, P, static_cast<double *>(x->d), static_cast<const double *>(b->d), x->n,
                                  opts->gcrk, opts->r_tol, opts->a_tol, opts->max_iter, opts->monitor_convergence, g)
               : fgmres_t<float>(A, P, static_cast<float *>(x->d), static_cast<const float *>(b->d), x->n,
                                 opts->gcrk, opts->r_tol, opts->a_tol, opts->max_iter, opts->monitor_convergence, g);
  if (rc) return rc;
  if (res) {
    res->status = g.status;
    res->iterations = g.cycles * opts->gcrk;
    res->outer_iterations = g.cycles;
    res->r_nrm_0 = g.r0;
    res->r_nrm = g.rn;
    res->true_rel_res = g.r0 > 0 ? g.rn / g.r0 : 0.0;
  }
  if (g.status == LFRIC_NOT_CONVERGED && !opts->fail_on_non_converged) return LFRIC_OK;
  if (g.status != LFRIC_OK) set_error(g.status == LFRIC_NAN ? "NaN in FGMRES residual" : "FGMRES did not converge");
  return g.status;
}

int lfric_ir_gmres_solve(const lfric_linop *A64, const lfric_linop *A32, const lfric_prec *P32, lfric_vec *x,
                         const lfric_vec *b, const lfric_solver_opts *opts, lfric_solver_result *res) {
  if (int rc = require_init()) return rc;
  LF_CHECK(A64 && A32 && x && b && opts, LFRIC_E_ARG, "null argument");
  LF_CHECK(A64->precision == LFRIC_R64 && A32->precision == LFRIC_R32 && A64->space == A32->space, LFRIC_E_ARG,
           "refinement needs FP64 and FP32 operators on one space");
  LF_CHECK(x->fs == A64->space && b->fs == A64->space && x->precision == LFRIC_R64 && b->precision == LFRIC_R64,
           LFRIC_E_ARG, "x and b must be FP64 vectors on the operator space");
  const long long n = x->n;
  void *r64, *b32, *d32;
  if (int rc = gm_ws(5, 8 * static_cast<size_t>(n) + 8 * static_cast<size_t>(n), &r64)) return rc;
  b32 = static_cast<char *>(r64) + 8 * static_cast<size_t>(n);
  d32 = static_cast<char *>(b32) + 4 * static_cast<size_t>(n);
  if (int rc = gm_init()) return rc;
  auto residual = [&](double *nrm) -> int {
    if (int rc = linop_apply(A64, x->d, r64)) return rc;
    if (int rc = launch_cg_residual(LFRIC_R64, r64, b->d, n)) return rc;
    if (int rc = dot_into<double>(static_cast<double *>(r64), static_cast<double *>(r64), n, &g_gstate->rnorm2, true))
      return rc;
    return read_scalar(&g_gstate->rnorm2, nrm);
  };
  double nr0, nr;
  if (int rc = residual(&nr0)) return rc;
  nr = nr0;
  int status = LFRIC_NOT_CONVERGED, cycles = 0, outer = 0, iterations = 0;
  if (!(nr0 == nr0)) status = LFRIC_NAN;
  else if (nr0 < 2.220446049250313e-16) status = LFRIC_OK;
  while (status == LFRIC_NOT_CONVERGED && outer < opts->max_outer && cycles < opts->max_iter) {
    if (nr / nr0 < opts->r_tol || nr <= opts->a_tol) {
      status = LFRIC_OK;
      break;
    }
    if (int rc = launch_convert(LFRIC_R32, b32, LFRIC_R64, r64, n, 1.0 / nr)) return rc;
    if (int rc = launch_set(LFRIC_R32, d32, n, 0.0)) return rc;
    lfric_solver_opts in = *opts;
    in.r_tol = std::max(opts->inner_r_tol, 0.5 * opts->r_tol * nr0 / nr);
    in.a_tol = 0.0;
    in.monitor_convergence = 1;
    in.max_iter = opts->max_iter - cycles;
    GmresOut g;
    if (int rc = gmres_common(A32, P32, d32, b32, n, LFRIC_R32, &in, g, true)) return rc;
    cycles += g.cycles;
    iterations += g.iterations;
    ++outer;
    if (g.status == LFRIC_NAN) {
      status = LFRIC_NAN;
      break;
    }
    if (int rc = launch_axpy_mixed(static_cast<double *>(x->d), nr, LFRIC_R32, d32, n)) return rc;
    if (int rc = residual(&nr)) return rc;
    if (!(nr == nr)) status = LFRIC_NAN;
  }
  if (status == LFRIC_NOT_CONVERGED && (nr / nr0 < opts->r_tol || nr <= opts->a_tol)) status = LFRIC_OK;
  if (res) {
    res->status = status;
    res->iterations = iterations;
    res->outer_iterations = outer;
    res->r_nrm_0 = nr0;
    res->r_nrm = nr;
    res->true_rel_res = nr0 > 0 ? nr / nr0 : 0.0;
  }
  if (status == LFRIC_NOT_CONVERGED && !opts->fail_on_non_converged) return LFRIC_OK;
  if (status != LFRIC_OK) set_error(status == LFRIC_NAN ? "NaN in refinement residual" : "refinement did not converge");
  return status;
}

}  // extern "C"
