// Batched solves of many independent LMA systems, each with several right-hand sides:
// FP32 (r_solver) CG inside FP64 iterative refinement for every (problem, RHS) at once.
//
// One inner CG iteration of the whole batch is a fixed chain of kernels whatever the number
// of problems: a persistent apply kernel that takes (problem, column block) tiles from a
// work counter (gather form of the LMA apply, so tiles of one problem have no colour order
// between them), then tile-flat vector kernels whose partial sums are finished per problem
// and RHS in tile order (deterministic).  The chain is captured once per active set in a HIP
// graph and replayed; the host looks at the per-RHS flags every round (check_every
// replays), runs the FP64 refinement step of the problems whose inner solve has converged,
// retires finished problems and admits new ones (continuous batching).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "device_common.h"
#include "internal.h"
#include "kernels.h"
#include "colour_apply.h"
#include "mfma_apply.h"
#include "sym27_apply.h"

namespace lfric {

namespace {

constexpr int kMaxR = 8;     // right-hand sides per problem (1, 2, 4 or 8)
constexpr int kMaxC = 8;     // colours of a problem whose apply runs on the matrix cores
constexpr int kVecTile = 4096;  // flat entries per vector tile
constexpr int kVecBlock = 256;

// device descriptor of one active problem
struct ProbDev {
  const float *A;
  unsigned long long gs;
  const int *ebase, *map2, *goff, *gent;
  const int2 *gcol;
  const float *diag;
  float *d, *r, *p, *q;
  long long n;      // undf * nrhs
  int ncols, nl, ndf, nrhs;
  int apply_tile0, vec_tile0, vec_ntiles, scal0;
  int8_t partner[32];
  // colour-ordered matrix-core apply (lowest-order W0): per colour, first tile, first cell
  // position and number of cell positions
  const int *map1;
  const uint32_t *first;
  int ctile0[kMaxC], cpos0[kMaxC], cnpos[kMaxC];
  // row-gathered copy (lowest-order W0): gather tiles over it
  const float *rows;
  unsigned long long gsr;
  // symmetric-packed order-1 copy: phase-1 unit tiles in colour slot 0, phase-2 gather tiles;
  // zs holds the per-unit row sums between the two
  const float *sym;
  unsigned long long gss;
  const int *sympos;
  const int4 *symzu;
  float *zs;
  long long units;
};

// per (problem, RHS) inner CG state
struct RhsDev {
  double rz, alpha, beta, r0, rn, tol;
  int done, iters;
};

__device__ inline int find_problem(const int *tile0, int nprob, int tile) {
  int lo = 0, hi = nprob;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile0[mid] <= tile)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// thread (column, level) of one problem, NR right-hand sides (see k_lma_gather_mrhs)
template <int N2, int NR>
__device__ inline void gather_point(const ProbDev &d, long long t) {
  const int H = d.nl + 1;
  if (t >= static_cast<long long>(d.ncols) * H) return;
  const int col = static_cast<int>(t / H);
  const int m = static_cast<int>(t - static_cast<long long>(col) * H);
  const int2 ci = d.gcol[col];
  if (m == d.nl && !ci.y) return;
  const bool has_m = m < d.nl, has_p = m > 0;
  typedef float VR __attribute__((ext_vector_type(NR)));
  VR acc = {};
  auto row = [&](long long e, int r, const int (&b)[N2], int lev) {
    for_row<float, N2>(d.A, d.gs, e, r, [&](int j, float a) {
      const VR xv = *reinterpret_cast<const VR *>(d.p + (static_cast<size_t>(b[j]) + lev) * NR);
#pragma unroll
      for (int k = 0; k < NR; ++k) acc[k] = fma(a, xv[k], acc[k]);
    });
  };
  const int q1 = d.goff[col + 1];
  for (int q = d.goff[col]; q < q1; ++q) {
    const int en = d.gent[q];
    const int pos = en >> 5, i = en & 31;
    const int pt = d.partner[i];
    const long long e = static_cast<long long>(d.ebase[pos]) + m;
    int b[N2];
#pragma unroll
    for (int j = 0; j < N2; ++j) b[j] = d.map2[static_cast<size_t>(pos) * N2 + j];
    if (pt >= 0 && has_p) row(e, pt, b, m - 1);
    if (has_m) row(e, i, b, m);
  }
  *reinterpret_cast<VR *>(d.q + (static_cast<size_t>(ci.x) + m) * NR) = acc;
}

// q = A32 p for the active problems with N2 dofs per cell and NR right-hand sides, whose tiles
// are [t_begin, t_end); persistent: blocks take tiles from *counter until none is left, so
// every wavefront exits.  One kernel per (N2, NR) group keeps each group's registers (and so
// its occupancy) its own.
template <int BS, int N2, int NR>
__global__ __launch_bounds__(BS) void k_batch_apply(const ProbDev *__restrict__ P, const int *__restrict__ tile0,
                                                    int nprob, int t_begin, int t_end, unsigned int *counter) {
  __shared__ int s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = t_begin + static_cast<int>(atomicAdd(counter, 1u));
    __syncthreads();
    // the tile is the same for the whole block: keep it (and the descriptor) in scalar registers
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    __syncthreads();
    if (tile >= t_end) return;
    const ProbDev &d = P[find_problem(tile0, nprob, tile)];
    gather_point<N2, NR>(d, static_cast<long long>(tile - d.apply_tile0) * BS + threadIdx.x);
  }
}

// q = A32 p in colour-ordered form on the matrix cores for one colour of the active lowest-order
// W0 problems of one nrhs, whose tiles (64 (cell, level) units of the colour each) are
// [t_begin, t_end); persistent like k_batch_apply.  The colours run as consecutive launches,
// so a colour's stores see the earlier colours' (no atomics).
template <int NR>
__global__ __launch_bounds__(kBlock) void k_batch_apply_mfma(const ProbDev *__restrict__ P,
                                                             const int *__restrict__ ctile0, int nprob, int colour,
                                                             int t_begin, int t_end, unsigned int *counter) {
  __shared__ int s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = t_begin + static_cast<int>(atomicAdd(counter, 1u));
    __syncthreads();
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    __syncthreads();
    if (tile >= t_end) return;
    const ProbDev &d = P[find_problem(ctile0, nprob, tile)];
    const int H = d.nl + 1;
    const long long u = static_cast<long long>(tile - d.ctile0[colour]) * (kBlock / 4) + (threadIdx.x >> 2);
    const bool valid = u < static_cast<long long>(d.cnpos[colour]) * H;
    const int lp = valid ? static_cast<int>(u / H) : 0;
    const int m = valid ? static_cast<int>(u - static_cast<long long>(lp) * H) : 0;
    const int pos = d.cpos0[colour] + lp;
    mfma_unit_w0<NR>(d.A, d.gs, static_cast<long long>(d.ebase[pos]) + m, d.map1 + static_cast<size_t>(pos) * 8,
                     d.map2 + static_cast<size_t>(pos) * 8, d.first[pos], d.p, d.q, m, d.nl, valid);
  }
}

// q = A32 p in colour-ordered form, one (cell, level) unit per thread, for one colour of the
// active order-1 W0 problems of one nrhs (tiles of 256 units); persistent like k_batch_apply.
template <int NR>
__global__ __launch_bounds__(kBlock) void k_batch_apply_colour27(const ProbDev *__restrict__ P,
                                                                 const int *__restrict__ ctile0, int nprob, int colour,
                                                                 int t_begin, int t_end, unsigned int *counter) {
  __shared__ int s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = t_begin + static_cast<int>(atomicAdd(counter, 1u));
    __syncthreads();
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    __syncthreads();
    if (tile >= t_end) return;
    const ProbDev &d = P[find_problem(ctile0, nprob, tile)];
    const int H = d.nl + 1;
    const long long u = static_cast<long long>(tile - d.ctile0[colour]) * kBlock + threadIdx.x;
    if (u < static_cast<long long>(d.cnpos[colour]) * H) {
      const int lp = static_cast<int>(u / H);
      const int m = static_cast<int>(u - static_cast<long long>(lp) * H);
      const int pos = d.cpos0[colour] + lp;
      colour_unit_mrhs<float, 27, 27, NR>(d.A, d.gs, static_cast<long long>(d.ebase[pos]) + m,
                                          d.map1 + static_cast<size_t>(pos) * 27, d.map2 + static_cast<size_t>(pos) * 27,
                                          d.first[pos], d.p, d.q, m, d.nl, RolesW0K1{});
    }
  }
}

// q = A32 p over the row-gathered copy for the active lowest-order W0 problems of one nrhs that
// have it, tiles of kBlock gather points [t_begin, t_end); persistent like k_batch_apply
template <int NR>
__global__ __launch_bounds__(kBlock) void k_batch_apply_rows(const ProbDev *__restrict__ P,
                                                             const int *__restrict__ tile0, int nprob, int t_begin,
                                                             int t_end, unsigned int *counter) {
  __shared__ int s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = t_begin + static_cast<int>(atomicAdd(counter, 1u));
    __syncthreads();
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    __syncthreads();
    if (tile >= t_end) return;
    const ProbDev &d = P[find_problem(tile0, nprob, tile)];
    rows_point_w0_mrhs<float, NR>(d.rows, d.gsr, d.p, d.q, d.map2, d.gcol, d.goff, d.gent, d.ncols, d.nl,
                                  static_cast<long long>(tile - d.apply_tile0) * kBlock + threadIdx.x);
  }
}

// phase 1 of the symmetric-packed order-1 apply (sym27_apply.h) for the active problems of one
// nrhs that have the copy: tiles of kBlock threads (NR/NRL per unit, whole wavefronts) in
// colour slot 0
template <int NR>
__global__ __launch_bounds__(kBlock) void k_batch_sym27_units(const ProbDev *__restrict__ P,
                                                              const int *__restrict__ ctile0, int nprob, int t_begin,
                                                              int t_end, unsigned int *counter) {
  constexpr int NRL = s27_nrl<float, NR>();
  __shared__ int s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = t_begin + static_cast<int>(atomicAdd(counter, 1u));
    __syncthreads();
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    __syncthreads();
    if (tile >= t_end) return;
    const ProbDev &d = P[find_problem(ctile0, nprob, tile)];
    sym27_unit<float, NR, NRL>(d.sym, d.gss, d.p, d.map2, d.sympos, d.zs, d.units, d.nl,
                               static_cast<long long>(tile - d.ctile0[0]) * kBlock + threadIdx.x);
  }
}

// phase 2: q at each gather point from the row sums, tiles of kBlock * kSymPts points (a thread
// takes kSymPts points kBlock apart: a point is a few loads, and one tile per kBlock points
// would make the work counter's atomics the bottleneck)
constexpr int kSymPts = 8;
template <int NR>
__global__ __launch_bounds__(kBlock) void k_batch_sym27_gather(const ProbDev *__restrict__ P,
                                                               const int *__restrict__ tile0, int nprob, int t_begin,
                                                               int t_end, unsigned int *counter) {
  typedef typename RVec<float, NR>::type VR;
  __shared__ int s_tile;
  for (;;) {
    if (threadIdx.x == 0) s_tile = t_begin + static_cast<int>(atomicAdd(counter, 1u));
    __syncthreads();
    const int tile = __builtin_amdgcn_readfirstlane(s_tile);
    __syncthreads();
    if (tile >= t_end) return;
    const ProbDev &d = P[find_problem(tile0, nprob, tile)];
    const long long t0 = static_cast<long long>(tile - d.apply_tile0) * kBlock * kSymPts + threadIdx.x;
    for (int j = 0; j < kSymPts; ++j) {
      VR acc;
      size_t dof = 0;
      if (sym27_gather_point<float, NR>(t0 + j * kBlock, d.zs, d.gcol, d.symzu, d.ncols, d.nl, acc, dof))
        *reinterpret_cast<VR *>(d.q + dof * NR) = acc;
    }
  }
}

// work counters: one per (apply group, colour)
constexpr int kMaxGroups = 16;
constexpr int kCounters = kMaxGroups * kMaxC;

__global__ void k_zero_counter(unsigned int *counter) {
  if (blockIdx.x == 0 && threadIdx.x < kCounters) counter[threadIdx.x] = 0u;
}

// per vector tile and RHS: sums of a.b (and c.d) over the tile's entries; thread t of the
// block always holds RHS t % nrhs, and the per-RHS sums are added in thread order
template <int NV>
__device__ inline void tile_sums(double (&v)[NV], int nrhs, double *out /* [NV][kMaxR] */) {
  __shared__ double sh[NV][kVecBlock];
#pragma unroll
  for (int k = 0; k < NV; ++k) sh[k][threadIdx.x] = v[k];
  __syncthreads();
  if (threadIdx.x < nrhs) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int u = threadIdx.x; u < kVecBlock; u += nrhs) s += sh[k][u];
      out[k * kMaxR + threadIdx.x] = s;
    }
  }
}

__device__ inline void vec_tile(const ProbDev *P, const int *vtile0, int nprob, const ProbDev *&d, long long &f0,
                                long long &f1) {
  const int tile = blockIdx.x;
  d = &P[find_problem(vtile0, nprob, tile)];
  f0 = static_cast<long long>(tile - d->vec_tile0) * kVecTile;
  f1 = std::min<long long>(f0 + kVecTile, d->n);
}

// partials of p.q
__global__ __launch_bounds__(kVecBlock) void k_batch_pq(const ProbDev *__restrict__ P, const int *__restrict__ vtile0,
                                                        int nprob, double *__restrict__ partials) {
  const ProbDev *d;
  long long f0, f1;
  vec_tile(P, vtile0, nprob, d, f0, f1);
  double v[1] = {0.0};
  for (long long f = f0 + threadIdx.x; f < f1; f += kVecBlock) v[0] += static_cast<double>(d->p[f]) * d->q[f];
  tile_sums<1>(v, d->nrhs, partials + static_cast<size_t>(blockIdx.x) * 2 * kMaxR);
}

// Total of one partial-sum column (entry idx of rows t0 .. t0+nt-1, rows of 2*kMaxR doubles)
// by one block of kFinBlock threads in a fixed order: thread k adds rows k, k+kFinBlock, ...
// in turn, then a fixed pairwise tree over the threads.  The same for every launch, so a
// problem's sums do not depend on the other problems of the batch.
constexpr int kFinBlock = 256;
__device__ inline double fixed_total(const double *__restrict__ partials, int t0, int nt, int idx) {
  __shared__ double sh[kFinBlock];
  double s = 0.0;
  for (int t = threadIdx.x; t < nt; t += kFinBlock) s += partials[static_cast<size_t>(t0 + t) * 2 * kMaxR + idx];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = kFinBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double tot = sh[0];
  __syncthreads();
  return tot;
}

// alpha = rz / p.q per (problem, RHS), from the tile partials; one block per problem
__global__ __launch_bounds__(kFinBlock) void k_batch_alpha(const ProbDev *__restrict__ P, int nprob,
                                                           const double *__restrict__ partials, RhsDev *__restrict__ S) {
  const int pi = blockIdx.x;
  if (pi >= nprob) return;
  const ProbDev &d = P[pi];
  for (int r = 0; r < d.nrhs; ++r) {
    RhsDev &s = S[d.scal0 + r];
    if (s.done) continue;  // the same for every thread of the block
    const double pq = fixed_total(partials, d.vec_tile0, d.vec_ntiles, r);
    if (threadIdx.x == 0) {
      if (pq != pq || pq == 0.0)
        s.done = 2;  // NaN or breakdown: stop this RHS (the host reports it)
      else
        s.alpha = s.rz / pq;
    }
  }
}

// d += alpha p; r -= alpha q; partials of r.r and r.z with z = r / diag (z is formed again
// where it is used, in k_batch_aypx, rather than stored)
__global__ __launch_bounds__(kVecBlock) void k_batch_update(const ProbDev *__restrict__ P,
                                                            const int *__restrict__ vtile0, int nprob,
                                                            const RhsDev *__restrict__ S, double *__restrict__ partials) {
  const ProbDev *d;
  long long f0, f1;
  vec_tile(P, vtile0, nprob, d, f0, f1);
  const int nr = d->nrhs;
  const int r = threadIdx.x % nr;
  const RhsDev &s = S[d->scal0 + r];
  const bool live = !s.done;
  const float a = static_cast<float>(s.alpha), ma = static_cast<float>(-s.alpha);
  double v[2] = {0.0, 0.0};
  for (long long f = f0 + threadIdx.x; f < f1; f += kVecBlock) {
    float ri = d->r[f];
    if (live) {
      d->d[f] = d->d[f] + a * d->p[f];
      ri = ri + ma * d->q[f];
      d->r[f] = ri;
      const float zi = ri / d->diag[f / nr];
      v[0] += static_cast<double>(ri) * ri;
      v[1] += static_cast<double>(ri) * zi;
    }
  }
  tile_sums<2>(v, nr, partials + static_cast<size_t>(blockIdx.x) * 2 * kMaxR);
}

// convergence test and beta of one (problem, RHS) from its r.r and r.z
__device__ inline void beta_step(RhsDev &s, double rr, double rz, double a_tol, int max_iter) {
  s.iters += 1;
  const double rn = sqrt(rr);
  s.rn = rn;
  if (rn != rn) {
    s.done = 2;
    return;
  }
  if (rn / s.r0 <= s.tol || rn <= a_tol || s.iters >= max_iter) {
    s.done = 1;
    return;
  }
  s.beta = rz / s.rz;
  s.rz = rz;
}

// ||r||, convergence test and beta per (problem, RHS); one block per problem
__global__ __launch_bounds__(kFinBlock) void k_batch_beta(const ProbDev *__restrict__ P, int nprob,
                                                          const double *__restrict__ partials, RhsDev *__restrict__ S,
                                                          double a_tol, int max_iter) {
  const int pi = blockIdx.x;
  if (pi >= nprob) return;
  const ProbDev &d = P[pi];
  for (int r = 0; r < d.nrhs; ++r) {
    if (S[d.scal0 + r].done) continue;  // the same for every thread of the block
    const double rr = fixed_total(partials, d.vec_tile0, d.vec_ntiles, r);
    const double rz = fixed_total(partials, d.vec_tile0, d.vec_ntiles, kMaxR + r);
    if (threadIdx.x == 0) beta_step(S[d.scal0 + r], rr, rz, a_tol, max_iter);
  }
}


// p = z + beta p for the live RHS, z = r / diag
__global__ __launch_bounds__(kVecBlock) void k_batch_aypx(const ProbDev *__restrict__ P, const int *__restrict__ vtile0,
                                                          int nprob, const RhsDev *__restrict__ S) {
  const ProbDev *d;
  long long f0, f1;
  vec_tile(P, vtile0, nprob, d, f0, f1);
  const int r = threadIdx.x % d->nrhs;
  const RhsDev &s = S[d->scal0 + r];
  if (s.done) return;
  const float b = static_cast<float>(s.beta);
  const int nr = d->nrhs;
  for (long long f = f0 + threadIdx.x; f < f1; f += kVecBlock) d->p[f] = b * d->p[f] + d->r[f] / d->diag[f / nr];
}

// ---- refinement step kernels (one problem at a time) ---------------------------------------
// x64 += scale[r] * d32 (the scaled inner solution)
__global__ void k_ir_accumulate(double *__restrict__ x, const float *__restrict__ d, const double *__restrict__ scale,
                                long long n, int nr) {
  for (long long f = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; f < n;
       f += static_cast<long long>(gridDim.x) * blockDim.x)
    x[f] = x[f] + scale[f % nr] * static_cast<double>(d[f]);
}

// *flag = 1 if any entry of x is not zero (set to 0 beforehand)
__global__ __launch_bounds__(kVecBlock) void k_nonzero(const double *__restrict__ x, long long n, int *flag) {
  for (long long f = static_cast<long long>(blockIdx.x) * kVecBlock + threadIdx.x; f < n;
       f += static_cast<long long>(gridDim.x) * kVecBlock)
    if (x[f] != 0.0) *flag = 1;
}

// r = b - r (in place), per-block partial sums of r.r per RHS
__global__ __launch_bounds__(kVecBlock) void k_ir_residual(double *__restrict__ r, const double *__restrict__ b,
                                                           long long n, int nr, double *__restrict__ partials) {
  double v[1] = {0.0};
  for (long long f = static_cast<long long>(blockIdx.x) * kVecBlock + threadIdx.x; f < n;
       f += static_cast<long long>(gridDim.x) * kVecBlock) {
    const double ri = b[f] - r[f];
    r[f] = ri;
    v[0] += ri * ri;
  }
  tile_sums<1>(v, nr, partials + static_cast<size_t>(blockIdx.x) * 2 * kMaxR);
}

__global__ __launch_bounds__(kFinBlock) void k_ir_norms(const double *__restrict__ partials, int nblocks, int nr,
                                                        double *__restrict__ out) {
  for (int r = 0; r < nr; ++r) {
    const double s = fixed_total(partials, 0, nblocks, r);
    if (threadIdx.x == 0) out[r] = sqrt(s);
  }
}

// start a new inner solve on rhs = r64 / scale[r]: d = 0, r = rhs, z = r / diag, p = z,
// partials of r.r and r.z per RHS (done RHS get zeros)
__global__ __launch_bounds__(kVecBlock) void k_ir_start(const double *__restrict__ r64, const double *__restrict__ inv,
                                                        float *__restrict__ d, float *__restrict__ r,
                                                        float *__restrict__ p,
                                                        const float *__restrict__ diag, long long n, int nr,
                                                        double *__restrict__ partials) {
  double v[2] = {0.0, 0.0};
  for (long long f = static_cast<long long>(blockIdx.x) * kVecBlock + threadIdx.x; f < n;
       f += static_cast<long long>(gridDim.x) * kVecBlock) {
    const float ri = static_cast<float>(r64[f] * inv[f % nr]);
    const float zi = ri / diag[f / nr];
    d[f] = 0.f;
    r[f] = ri;
    p[f] = zi;
    v[0] += static_cast<double>(ri) * ri;
    v[1] += static_cast<double>(ri) * zi;
  }
  tile_sums<2>(v, nr, partials + static_cast<size_t>(blockIdx.x) * 2 * kMaxR);
}

__device__ inline void start_rhs(RhsDev &s, double rr, double rz, double tol, int live) {
  s.rz = rz;
  s.r0 = sqrt(rr);
  s.rn = s.r0;
  s.tol = tol;
  s.alpha = s.beta = 0.0;
  s.iters = 0;
  s.done = (live && s.r0 > 0.0) ? 0 : 1;
}

__global__ __launch_bounds__(kFinBlock) void k_ir_start_fin(const double *__restrict__ partials, int nblocks, int nr,
                                                            RhsDev *__restrict__ S, const double *__restrict__ tol,
                                                            const int *__restrict__ live) {
  for (int r = 0; r < nr; ++r) {
    const double rr = fixed_total(partials, 0, nblocks, r);
    const double rz = fixed_total(partials, 0, nblocks, kMaxR + r);
    if (threadIdx.x == 0) start_rhs(S[r], rr, rz, tol[r], live[r]);
  }
}

constexpr int kIrBlocks = 512;

// ---- host side ---------------------------------------------------------------------------------
struct Problem {
  int id = -1;
  const lfric_op *A64 = nullptr, *A32 = nullptr;
  const lfric_vec *diag = nullptr, *b = nullptr;
  lfric_vec *x = nullptr;
  int nrhs = 1;
  long long n = 0;
  int scal0 = 0;  // first of this problem's nrhs entries of the batch's RhsDev array
  float *d = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;  // z = r / diag is not stored
  double *r64 = nullptr;
  size_t bytes32 = 0, bytes64 = 0;
  float *zs = nullptr;  // per-unit row sums of the symmetric-packed order-1 apply
  size_t zs_bytes = 0;
  // refinement state per RHS
  std::vector<double> nr0, nr;
  std::vector<int> conv, inner_iters;
  int outer = 0, status = LFRIC_NOT_CONVERGED;
  bool started = false, finished = false, inner_running = false;
};

}  // namespace
}  // namespace lfric

struct lfric_batch {
  lfric_solver_opts opts{};
  std::vector<std::unique_ptr<lfric::Problem>> probs;  // every problem not yet removed
  int next_id = 0;
  // per-(problem, RHS) state; ranges of freed problems are reused by size
  lfric::RhsDev *d_state = nullptr, *h_state = nullptr;
  size_t cap_state = 0, used_state = 0;
  std::vector<std::pair<int, int>> free_ranges;  // (scal0, nrhs)
  // tables of the active set
  lfric::ProbDev *d_desc = nullptr;
  int *d_tile0 = nullptr, *d_vtile0 = nullptr, *d_ctile0 = nullptr;
  size_t cap_desc = 0, cap_tile0 = 0, cap_vtile0 = 0, cap_ctile0 = 0, cap_partials = 0;
  double *d_partials = nullptr, *d_scratch = nullptr, *h_scratch = nullptr;
  size_t scratch_slots = 0;  // problems whose refinement step the scratch holds at once
  unsigned int *d_counter = nullptr;
  std::vector<lfric::Problem *> active;
  int ntiles = 0, nvtiles = 0;
  // apply groups of equal (ndf, nrhs, form): gather tiles [begin, end) (forms 0, 3 and the
  // second phase of 4), colour tiles [cbegin[c], cend[c]) of each colour c of the colour-
  // ordered forms (1 matrix cores, 2 register-staged FMA), the first phase's unit tiles of form
  // 4 in [cbegin[0], cend[0])
  struct Group {
    int ndf, nrhs, form, begin, end;
    int cbegin[lfric::kMaxC], cend[lfric::kMaxC];
  };
  int np_tables = 0;  // problems in the tables (row length of the colour tile tables - 1)
  std::vector<Group> groups;
  bool dirty = true;
  // graph of one inner iteration of the active set
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int use_graph = 1;
  int use_colour = 2;  // apply forms (lfric_batch_set_apply): copies, then colour-ordered forms
  // launch shape of the apply (autotuned)
  int block = 256, grid_mult = 2, ncu = 256;
  long long replays = 0, iterations = 0;
};

namespace lfric {
namespace {

#define LAUNCHB(kernel, grid, block, ...) \
  hipLaunchKernelGGL((kernel), dim3(grid), dim3(block), 0, cur_stream(), __VA_ARGS__)

constexpr double kEps = 2.220446049250313e-16;

int ensure(void **p, size_t &cap, size_t need, size_t elem) {
  if (need <= cap && *p) return LFRIC_OK;
  dfree(*p, cap * elem);
  *p = nullptr;
  cap = 0;
  const size_t c = std::max<size_t>(need, 16) * 2;
  if (int rc = dmalloc(p, c * elem)) return rc;
  cap = c;
  return LFRIC_OK;
}

void drop_graph(lfric_batch &B) {
  if (B.exec) (void)hipGraphExecDestroy(B.exec);
  if (B.graph) (void)hipGraphDestroy(B.graph);
  B.exec = nullptr;
  B.graph = nullptr;
}

int apply_tiles(const Problem &p, int block) {
  const long long pts = static_cast<long long>(p.A32->to->gather_ncols) * (p.A32->to->nlayers + 1);
  return static_cast<int>((pts + block - 1) / block);
}

// the problem's apply form in the batch: 0 gather, 1 colour-ordered on the matrix cores, 2
// colour-ordered with register-staged x, 3 gather over the row-gathered copy, 4 the two phases
// of the symmetric-packed order-1 copy.  B.use_colour: 0 gather for every problem, 1 the
// colour-ordered forms where supported, 2 (default) the operator's copies where it has one,
// else as 1.
constexpr int kFormRows = 3, kFormSym27 = 4;

bool is_lowest_w0_fp32(const lfric_op *op) { return mfma_apply_supported(op); }

int apply_form(const lfric_batch &B, const Problem &p) {
  if (!B.use_colour) return 0;
  if (B.use_colour == 2) {
    if (p.A32->rows && is_lowest_w0_fp32(p.A32)) return kFormRows;
    if (p.A32->precision == LFRIC_R32 && sym27_ready(p.A32)) return kFormSym27;
  }
  if (p.A32->to->ncolours > kMaxC) return 0;
  return batch_colour_form(p.A32);
}

bool colour_form(int form) { return form == 1 || form == 2; }

// (cell, level) units per tile of a colour-ordered form
int units_per_tile(int form) { return form == 1 ? kBlock / 4 : kBlock; }

// threads of the symmetric-packed order-1 apply's first phase
long long sym27_threads(int nrhs, long long units) {
  switch (nrhs) {
    case 1: return s27_threads<float, 1>(units);
    case 2: return s27_threads<float, 2>(units);
    case 4: return s27_threads<float, 4>(units);
    default: return s27_threads<float, 8>(units);
  }
}

// state entries for a new problem (reusing a freed range of the same size), growing the array
// with a device copy when needed
int alloc_state(lfric_batch &B, int nr, int *scal0) {
  for (size_t i = 0; i < B.free_ranges.size(); ++i)
    if (B.free_ranges[i].second == nr) {
      *scal0 = B.free_ranges[i].first;
      B.free_ranges.erase(B.free_ranges.begin() + static_cast<long>(i));
      return LFRIC_OK;
    }
  const size_t need = B.used_state + nr;
  if (need > B.cap_state) {
    const size_t cap = std::max<size_t>(need * 2, 64);
    RhsDev *d = nullptr, *h = nullptr;
    if (int rc = dmalloc(reinterpret_cast<void **>(&d), cap * sizeof(RhsDev))) return rc;
    if (hipHostMalloc(reinterpret_cast<void **>(&h), cap * sizeof(RhsDev), hipHostMallocDefault) != hipSuccess) {
      dfree(d, cap * sizeof(RhsDev));
      return fail(LFRIC_E_NOMEM, "pinned host state");
    }
    if (B.used_state) {
      LF_HIP(hipMemcpyAsync(d, B.d_state, B.used_state * sizeof(RhsDev), hipMemcpyDeviceToDevice, cur_stream()));
      LF_HIP(hipStreamSynchronize(cur_stream()));
    }
    dfree(B.d_state, B.cap_state * sizeof(RhsDev));
    if (B.h_state) (void)hipHostFree(B.h_state);
    B.d_state = d;
    B.h_state = h;
    B.cap_state = cap;
    B.dirty = true;  // kernels hold the array's address
  }
  *scal0 = static_cast<int>(B.used_state);
  B.used_state = need;
  return LFRIC_OK;
}

// device tables of the problems whose inner solve runs; the graph is rebuilt on next use
int rebuild(lfric_batch &B) {
  B.active.clear();
  for (auto &u : B.probs)
    if (u->inner_running) B.active.push_back(u.get());
  for (Problem *p : B.active)
    if (apply_form(B, *p) == kFormSym27 && !p->zs) {
      const size_t need = sizeof(float) * s27_zalloc(p->A32->to->ncells, p->A32->to->nlayers, p->nrhs);
      if (int rc = dmalloc(reinterpret_cast<void **>(&p->zs), need)) return rc;
      p->zs_bytes = need;
    }
  // order-1 problems first (their own apply kernel), then the gather-form and matrix-core groups,
  // heaviest tiles first within each kind
  std::stable_sort(B.active.begin(), B.active.end(), [&B](const Problem *a, const Problem *b) {
    const int ka = a->A32->from->ndf, kb = b->A32->from->ndf;
    if (ka != kb) return ka > kb;
    const int fa = apply_form(B, *a), fb = apply_form(B, *b);
    if (fa != fb) return fa < fb;
    if (a->nrhs != b->nrhs) return a->nrhs > b->nrhs;
    return a->n > b->n;
  });
  const int np = static_cast<int>(B.active.size());
  std::vector<ProbDev> desc(np);
  std::vector<int> tile0(np + 1, 0), vtile0(np + 1, 0), ctile0(static_cast<size_t>(kMaxC) * (np + 1), 0);
  std::vector<int> mf(np, 0);
  for (int i = 0; i < np; ++i) {
    const Problem &p = *B.active[i];
    const FunctionSpace &to = *p.A32->to, &fr = *p.A32->from;
    ProbDev &d = desc[i];
    std::memset(&d, 0, sizeof(d));
    d.A = static_cast<const float *>(p.A32->d);
    d.gs = p.A32->gs;
    d.ebase = to.d_ebase;
    d.map2 = fr.d_map;
    d.goff = to.d_goff;
    d.gent = to.d_gent;
    d.gcol = to.d_gcol;
    d.diag = static_cast<const float *>(p.diag->d);
    d.d = p.d;
    d.r = p.r;
    d.p = p.p;
    d.q = p.q;
    d.n = p.n;
    d.ncols = to.gather_ncols;
    d.nl = to.nlayers;
    d.ndf = fr.ndf;
    d.nrhs = p.nrhs;
    for (int k = 0; k < 32; ++k) d.partner[k] = static_cast<int8_t>(k < to.ndf ? to.partner[k] : -1);
    mf[i] = apply_form(B, p);
    d.map1 = to.d_map;
    d.first = to.d_first;
    if (mf[i] == kFormRows) {
      d.rows = static_cast<const float *>(p.A32->rows);
      d.gsr = p.A32->gsr;
    }
    if (mf[i] == kFormSym27) {
      d.sym = static_cast<const float *>(p.A32->sym);
      d.gss = p.A32->gss;
      d.sympos = p.A32->d_sympos;
      d.symzu = p.A32->d_symzu;
      d.zs = p.zs;
      d.units = static_cast<long long>(to.ncells) * to.nlayers;
    }
    for (int c = 0; c < kMaxC; ++c) {
      int *row = ctile0.data() + static_cast<size_t>(c) * (np + 1);
      const bool has = colour_form(mf[i]) && c < to.ncolours;
      d.cpos0[c] = has ? to.colour_start[c] : 0;
      d.cnpos[c] = has ? to.colour_start[c + 1] - to.colour_start[c] : 0;
      d.ctile0[c] = row[i];
      long long threads = static_cast<long long>(d.cnpos[c]) * (to.nlayers + 1);
      int upt = units_per_tile(mf[i]);
      if (mf[i] == kFormSym27 && c == 0) {  // phase-1 tiles (whole wavefronts)
        threads = sym27_threads(p.nrhs, d.units);
        upt = kBlock;
      }
      row[i + 1] = row[i] + static_cast<int>((threads + upt - 1) / upt);
    }
    d.apply_tile0 = tile0[i];
    const int gtile = mf[i] == kFormSym27 ? kBlock * kSymPts : mf[i] == kFormRows ? kBlock : B.block;
    tile0[i + 1] = tile0[i] + (colour_form(mf[i]) ? 0 : apply_tiles(p, gtile));
    d.vec_tile0 = vtile0[i];
    d.vec_ntiles = static_cast<int>((p.n + kVecTile - 1) / kVecTile);
    vtile0[i + 1] = vtile0[i] + d.vec_ntiles;
    d.scal0 = p.scal0;
  }
  B.ntiles = tile0[np];
  B.nvtiles = vtile0[np];
  B.groups.clear();
  for (int i = 0; i < np; ++i) {
    if (B.groups.empty() || B.groups.back().ndf != desc[i].ndf || B.groups.back().nrhs != desc[i].nrhs ||
        B.groups.back().form != mf[i]) {
      lfric_batch::Group g{};
      g.ndf = desc[i].ndf;
      g.nrhs = desc[i].nrhs;
      g.form = mf[i];
      g.begin = g.end = tile0[i];
      for (int c = 0; c < kMaxC; ++c) g.cbegin[c] = g.cend[c] = ctile0[static_cast<size_t>(c) * (np + 1) + i];
      B.groups.push_back(g);
    }
    B.groups.back().end = tile0[i + 1];
    for (int c = 0; c < kMaxC; ++c) B.groups.back().cend[c] = ctile0[static_cast<size_t>(c) * (np + 1) + i + 1];
  }
  LF_CHECK(static_cast<int>(B.groups.size()) <= kMaxGroups, LFRIC_E_UNSUPPORTED, "too many apply groups");
  B.np_tables = np;
  if (int rc = ensure(reinterpret_cast<void **>(&B.d_desc), B.cap_desc, np + 1, sizeof(ProbDev))) return rc;
  if (int rc = ensure(reinterpret_cast<void **>(&B.d_tile0), B.cap_tile0, np + 1, sizeof(int))) return rc;
  if (int rc = ensure(reinterpret_cast<void **>(&B.d_vtile0), B.cap_vtile0, np + 1, sizeof(int))) return rc;
  if (int rc = ensure(reinterpret_cast<void **>(&B.d_ctile0), B.cap_ctile0, ctile0.size(), sizeof(int))) return rc;
  if (int rc = ensure(reinterpret_cast<void **>(&B.d_partials), B.cap_partials,
                      static_cast<size_t>(std::max(B.nvtiles, kIrBlocks)) * 2 * kMaxR, sizeof(double)))
    return rc;
  if (np) {
    LF_HIP(hipMemcpyAsync(B.d_desc, desc.data(), sizeof(ProbDev) * np, hipMemcpyHostToDevice, cur_stream()));
    LF_HIP(hipMemcpyAsync(B.d_tile0, tile0.data(), sizeof(int) * (np + 1), hipMemcpyHostToDevice, cur_stream()));
    LF_HIP(hipMemcpyAsync(B.d_vtile0, vtile0.data(), sizeof(int) * (np + 1), hipMemcpyHostToDevice, cur_stream()));
    LF_HIP(hipMemcpyAsync(B.d_ctile0, ctile0.data(), sizeof(int) * ctile0.size(), hipMemcpyHostToDevice, cur_stream()));
  }
  LF_HIP(hipStreamSynchronize(cur_stream()));
  drop_graph(B);
  B.dirty = false;
  return LFRIC_OK;
}

template <int BS, int N2>
void launch_group(lfric_batch &B, int g, int cap) {
  const auto &G = B.groups[g];
  const int np = static_cast<int>(B.active.size());
  const int grid = std::min(G.end - G.begin, cap);
  unsigned int *cnt = B.d_counter + g * kMaxC;
  switch (G.nrhs) {
    case 1: LAUNCHB((k_batch_apply<BS, N2, 1>), grid, BS, B.d_desc, B.d_tile0, np, G.begin, G.end, cnt); break;
    case 2: LAUNCHB((k_batch_apply<BS, N2, 2>), grid, BS, B.d_desc, B.d_tile0, np, G.begin, G.end, cnt); break;
    case 4: LAUNCHB((k_batch_apply<BS, N2, 4>), grid, BS, B.d_desc, B.d_tile0, np, G.begin, G.end, cnt); break;
    default: LAUNCHB((k_batch_apply<BS, N2, 8>), grid, BS, B.d_desc, B.d_tile0, np, G.begin, G.end, cnt); break;
  }
}

// the colours of a colour-ordered group, one persistent launch each
void launch_group_colour(lfric_batch &B, int g, int grid_mult) {
  const auto &G = B.groups[g];
  const int np = B.np_tables;
  const int cap = grid_mult * B.ncu * (1024 / kBlock);
  for (int c = 0; c < kMaxC; ++c) {
    if (G.cend[c] == G.cbegin[c]) continue;
    const int grid = std::min(G.cend[c] - G.cbegin[c], cap);
    const int *tab = B.d_ctile0 + static_cast<size_t>(c) * (np + 1);
    unsigned int *cnt = B.d_counter + g * kMaxC + c;
#define COLOUR_LAUNCH(K) LAUNCHB(K, grid, kBlock, B.d_desc, tab, np, c, G.cbegin[c], G.cend[c], cnt)
    if (G.form == 1) {
      switch (G.nrhs) {
        case 1: COLOUR_LAUNCH(k_batch_apply_mfma<1>); break;
        case 2: COLOUR_LAUNCH(k_batch_apply_mfma<2>); break;
        case 4: COLOUR_LAUNCH(k_batch_apply_mfma<4>); break;
        default: COLOUR_LAUNCH(k_batch_apply_mfma<8>); break;
      }
    } else {
      switch (G.nrhs) {
        case 1: COLOUR_LAUNCH(k_batch_apply_colour27<1>); break;
        case 2: COLOUR_LAUNCH(k_batch_apply_colour27<2>); break;
        case 4: COLOUR_LAUNCH(k_batch_apply_colour27<4>); break;
        default: COLOUR_LAUNCH(k_batch_apply_colour27<8>); break;
      }
    }
#undef COLOUR_LAUNCH
  }
}

// the gather tiles of a group over its row-gathered copies
void launch_group_rows(lfric_batch &B, int g, int cap) {
  const auto &G = B.groups[g];
  if (G.end == G.begin) return;
  const int np = B.np_tables, grid = std::min(G.end - G.begin, cap);
  unsigned int *cnt = B.d_counter + g * kMaxC;
#define ROWS_LAUNCH(K) LAUNCHB(K, grid, kBlock, B.d_desc, B.d_tile0, np, G.begin, G.end, cnt)
  switch (G.nrhs) {
    case 1: ROWS_LAUNCH(k_batch_apply_rows<1>); break;
    case 2: ROWS_LAUNCH(k_batch_apply_rows<2>); break;
    case 4: ROWS_LAUNCH(k_batch_apply_rows<4>); break;
    default: ROWS_LAUNCH(k_batch_apply_rows<8>); break;
  }
#undef ROWS_LAUNCH
}

// one phase of a symmetric-packed order-1 group: 1 the unit tiles (colour slot 0 of the
// colour tile table), 2 the gather tiles
void launch_group_sym27(lfric_batch &B, int g, int phase, int cap) {
  const auto &G = B.groups[g];
  const int np = B.np_tables;
  const int b0 = phase == 1 ? G.cbegin[0] : G.begin, b1 = phase == 1 ? G.cend[0] : G.end;
  if (b1 == b0) return;
  const int grid = std::min(b1 - b0, cap);
  unsigned int *cnt = B.d_counter + g * kMaxC + (phase == 1 ? 0 : 1);
  const int *tab = phase == 1 ? B.d_ctile0 : B.d_tile0;
#define SYM_LAUNCH(K) LAUNCHB(K, grid, kBlock, B.d_desc, tab, np, b0, b1, cnt)
  if (phase == 1) {
    switch (G.nrhs) {
      case 1: SYM_LAUNCH(k_batch_sym27_units<1>); break;
      case 2: SYM_LAUNCH(k_batch_sym27_units<2>); break;
      case 4: SYM_LAUNCH(k_batch_sym27_units<4>); break;
      default: SYM_LAUNCH(k_batch_sym27_units<8>); break;
    }
  } else {
    switch (G.nrhs) {
      case 1: SYM_LAUNCH(k_batch_sym27_gather<1>); break;
      case 2: SYM_LAUNCH(k_batch_sym27_gather<2>); break;
      case 4: SYM_LAUNCH(k_batch_sym27_gather<4>); break;
      default: SYM_LAUNCH(k_batch_sym27_gather<8>); break;
    }
  }
#undef SYM_LAUNCH
}

template <int BS>
void launch_apply_bs(lfric_batch &B, int grid_mult) {
  const int cap = grid_mult * B.ncu * (1024 / BS), capk = grid_mult * B.ncu * (1024 / kBlock);
  const int ng = static_cast<int>(B.groups.size());
  for (int g = 0; g < ng; ++g)
    if (B.groups[g].form == kFormSym27) launch_group_sym27(B, g, 1, capk);
  for (int g = 0; g < ng; ++g) {
    if (B.groups[g].form == kFormSym27) continue;
    if (B.groups[g].form == kFormRows) {
      launch_group_rows(B, g, capk);
      continue;
    }
    if (B.groups[g].form) {
      launch_group_colour(B, g, grid_mult);
      continue;
    }
    if (B.groups[g].end == B.groups[g].begin) continue;
    if (B.groups[g].ndf == 27)
      launch_group<BS, 27>(B, g, cap);
    else
      launch_group<BS, 8>(B, g, cap);
  }
  for (int g = 0; g < ng; ++g)
    if (B.groups[g].form == kFormSym27) launch_group_sym27(B, g, 2, capk);
}

void launch_apply(lfric_batch &B, int block, int grid_mult) {
  LAUNCHB(k_zero_counter, 1, kCounters, B.d_counter);
  switch (block) {
    case 64: launch_apply_bs<64>(B, grid_mult); break;
    case 128: launch_apply_bs<128>(B, grid_mult); break;
    default: launch_apply_bs<256>(B, grid_mult); break;
  }
}

// one inner CG iteration of every active problem (the reference's cg_solve loop body,
// sci_iterative_solver_mod.f90:545-575, per RHS)
int enqueue_iteration(lfric_batch &B) {
  const int np = static_cast<int>(B.active.size());
  launch_apply(B, B.block, B.grid_mult);                                                       // q = A p
  LAUNCHB(k_batch_pq, B.nvtiles, kVecBlock, B.d_desc, B.d_vtile0, np, B.d_partials);           // p.q
  LAUNCHB(k_batch_alpha, np, kFinBlock, B.d_desc, np, B.d_partials, B.d_state);               // alpha
  LAUNCHB(k_batch_update, B.nvtiles, kVecBlock, B.d_desc, B.d_vtile0, np, B.d_state, B.d_partials);
  LAUNCHB(k_batch_beta, np, kFinBlock, B.d_desc, np, B.d_partials, B.d_state, 0.0, B.opts.max_iter);  // test, beta
  LAUNCHB(k_batch_aypx, B.nvtiles, kVecBlock, B.d_desc, B.d_vtile0, np, B.d_state);            // p = z + beta p
  LF_HIP(hipGetLastError());
  return LFRIC_OK;
}

int run_iterations(lfric_batch &B, int k) {
  if (B.active.empty()) return LFRIC_OK;
  if (B.use_graph) {
    if (!B.exec) {
      LF_HIP(hipStreamBeginCapture(cur_stream(), hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_iteration(B);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(cur_stream(), &g);
      if (rc) return rc;
      if (e != hipSuccess) return fail(LFRIC_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      B.graph = g;
      LF_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
    }
    for (int i = 0; i < k; ++i) LF_HIP(hipGraphLaunch(B.exec, cur_stream()));
    B.replays += k;
  } else {
    for (int i = 0; i < k; ++i)
      if (int rc = enqueue_iteration(B)) return rc;
  }
  B.iterations += k;
  return LFRIC_OK;
}

// FP64 refinement step of one problem (the outer loop of the mixed-precision CG): fold the
// scaled inner solution into x, take b - A64 x in FP64, test each RHS, and start the next
// inner solve on r / ||r|| for the RHS that still need one
// per problem of a refinement pass: scale, 1/||r||, inner tolerance, live flags (as ints), the
// residual norms and a nonzero-x flag, kMaxR each, in the batch's device and pinned host scratch
constexpr int kSlot = 6 * kMaxR;

int ensure_slots(lfric_batch &B, size_t n) {
  if (n <= B.scratch_slots) return LFRIC_OK;
  const size_t slots = std::max<size_t>(n, 2 * B.scratch_slots);
  LF_HIP(hipStreamSynchronize(cur_stream()));
  dfree(B.d_scratch, sizeof(double) * kSlot * B.scratch_slots);
  if (B.h_scratch) (void)hipHostFree(B.h_scratch);
  B.d_scratch = B.h_scratch = nullptr;
  B.scratch_slots = 0;
  if (int rc = dmalloc(reinterpret_cast<void **>(&B.d_scratch), sizeof(double) * kSlot * slots)) return rc;
  if (hipHostMalloc(reinterpret_cast<void **>(&B.h_scratch), sizeof(double) * kSlot * slots, hipHostMallocDefault) !=
      hipSuccess) {
    dfree(B.d_scratch, sizeof(double) * kSlot * slots);
    B.d_scratch = nullptr;
    return fail(LFRIC_E_NOMEM, "pinned host scratch");
  }
  B.scratch_slots = slots;
  return LFRIC_OK;
}

// The FP64 refinement step of the problems in ps (each the same as on its own): add the scaled
// inner solution into x, take b - A64 x in FP64 and its norm per RHS, then either retire the
// problem or restart its inner solve on r/||r||.  Every problem's kernels are queued before the
// host waits, twice in all.
int refine_all(lfric_batch &B, const std::vector<Problem *> &ps) {
  if (ps.empty()) return LFRIC_OK;
  if (int rc = ensure_slots(B, ps.size())) return rc;
  // a problem's first step with x = 0 (the usual start) has A64 x = 0: its FP64 apply is
  // replaced by zeros (one more host wait for all such problems together)
  bool first = false;
  for (size_t i = 0; i < ps.size(); ++i) {
    Problem &p = *ps[i];
    if (p.started) continue;
    int *flag = reinterpret_cast<int *>(B.d_scratch + i * kSlot + 5 * kMaxR);
    LF_HIP(hipMemsetAsync(flag, 0, sizeof(int), cur_stream()));
    LAUNCHB(k_nonzero, grid_for(p.n, kVecBlock, 2048), kVecBlock, static_cast<const double *>(p.x->d), p.n, flag);
    LF_HIP(hipMemcpyAsync(B.h_scratch + i * kSlot + 5 * kMaxR, flag, sizeof(int), hipMemcpyDeviceToHost,
                          cur_stream()));
    first = true;
  }
  if (first) LF_HIP(hipStreamSynchronize(cur_stream()));
  for (size_t i = 0; i < ps.size(); ++i) {
    Problem &p = *ps[i];
    double *d = B.d_scratch + i * kSlot, *h = B.h_scratch + i * kSlot;
    const int nr = p.nrhs;
    const long long n = p.n;
    if (p.started) {
      for (int r = 0; r < nr; ++r) h[r] = p.conv[r] ? 0.0 : p.nr[r];
      LF_HIP(hipMemcpyAsync(d, h, sizeof(double) * nr, hipMemcpyHostToDevice, cur_stream()));
      LAUNCHB(k_ir_accumulate, grid_for(n, kBlock, 8192), kBlock, static_cast<double *>(p.x->d), p.d, d, n, nr);
    }
    if (!p.started && *reinterpret_cast<const int *>(h + 5 * kMaxR) == 0) {
      LF_HIP(hipMemsetAsync(p.r64, 0, sizeof(double) * static_cast<size_t>(n), cur_stream()));
    } else if (int rc = launch_lma_apply_mrhs(p.A64, p.x->d, p.r64, nr, LFRIC_MRHS_AUTO)) {
      return rc;
    }
    LAUNCHB(k_ir_residual, kIrBlocks, kVecBlock, p.r64, static_cast<const double *>(p.b->d), n, nr, B.d_partials);
    LAUNCHB(k_ir_norms, 1, kFinBlock, B.d_partials, kIrBlocks, nr, d + 4 * kMaxR);
    LF_HIP(hipMemcpyAsync(h + 4 * kMaxR, d + 4 * kMaxR, sizeof(double) * nr, hipMemcpyDeviceToHost, cur_stream()));
  }
  LF_HIP(hipStreamSynchronize(cur_stream()));
  for (size_t i = 0; i < ps.size(); ++i) {
    Problem &p = *ps[i];
    double *d = B.d_scratch + i * kSlot, *h = B.h_scratch + i * kSlot;
    const int nr = p.nrhs;
    bool all = true, nan = false;
    for (int r = 0; r < nr; ++r) {
      const double v = h[4 * kMaxR + r];
      if (!p.started) p.nr0[r] = v;
      if (!p.conv[r]) p.nr[r] = v;
      if (v != v) nan = true;
      if (!p.conv[r] && (v <= B.opts.r_tol * p.nr0[r] || v <= B.opts.a_tol || p.nr0[r] < kEps)) p.conv[r] = 1;
      all = all && p.conv[r];
    }
    if (p.started) ++p.outer;
    p.started = true;
    if (nan || all || p.outer >= B.opts.max_outer) {
      p.status = nan ? LFRIC_NAN : all ? LFRIC_OK : LFRIC_NOT_CONVERGED;
      p.finished = true;
      if (p.inner_running) B.dirty = true;
      p.inner_running = false;
      continue;
    }
    double *inv = d + kMaxR, *tol = d + 2 * kMaxR;
    int *live = reinterpret_cast<int *>(d + 3 * kMaxR);
    for (int r = 0; r < 3 * kMaxR; ++r) h[kMaxR + r] = 0.0;
    for (int r = 0; r < nr; ++r) {
      h[kMaxR + r] = p.conv[r] ? 0.0 : 1.0 / p.nr[r];
      h[2 * kMaxR + r] = std::max(B.opts.inner_r_tol, 0.5 * B.opts.r_tol * p.nr0[r] / p.nr[r]);
      reinterpret_cast<int *>(h + 3 * kMaxR)[r] = p.conv[r] ? 0 : 1;
    }
    LF_HIP(hipMemcpyAsync(inv, h + kMaxR, sizeof(double) * 3 * kMaxR, hipMemcpyHostToDevice, cur_stream()));
    LAUNCHB(k_ir_start, kIrBlocks, kVecBlock, p.r64, inv, p.d, p.r, p.p, static_cast<const float *>(p.diag->d),
            p.n, nr, B.d_partials);
    LAUNCHB(k_ir_start_fin, 1, kFinBlock, B.d_partials, kIrBlocks, nr, B.d_state + p.scal0, tol, live);
    if (!p.inner_running) B.dirty = true;
    p.inner_running = true;
  }
  LF_HIP(hipGetLastError());
  LF_HIP(hipStreamSynchronize(cur_stream()));
  return LFRIC_OK;
}

void free_problem(lfric_batch &B, Problem &p) {
  dfree(p.d, p.bytes32);
  dfree(p.r64, p.bytes64);
  dfree(p.zs, p.zs_bytes);
  p.d = nullptr;
  p.r64 = nullptr;
  p.zs = nullptr;
  p.zs_bytes = 0;
  if (p.scal0 >= 0) B.free_ranges.push_back({p.scal0, p.nrhs});
  p.scal0 = -1;
}

Problem *find(lfric_batch &B, int id) {
  for (auto &u : B.probs)
    if (u->id == id) return u.get();
  return nullptr;
}

int run_rounds(lfric_batch &B, int rounds, int *nactive) {
  const int k = std::max(1, B.opts.check_every);
  for (int round = 0; round < rounds; ++round) {
    std::vector<Problem *> due;
    for (auto &u : B.probs)  // problems added since the last round: first residual
      if (!u->started && !u->finished) due.push_back(u.get());
    if (int rc = refine_all(B, due)) return rc;
    if (B.dirty)
      if (int rc = rebuild(B)) return rc;
    if (B.active.empty()) break;
    if (int rc = run_iterations(B, k)) return rc;
    LF_HIP(hipMemcpyAsync(B.h_state, B.d_state, sizeof(RhsDev) * B.used_state, hipMemcpyDeviceToHost,
                          cur_stream()));
    LF_HIP(hipStreamSynchronize(cur_stream()));
    due.clear();
    for (Problem *p : B.active) {
      bool ended = true;
      for (int r = 0; r < p->nrhs; ++r)
        if (!p->conv[r] && !B.h_state[p->scal0 + r].done) ended = false;
      if (!ended) continue;
      for (int r = 0; r < p->nrhs; ++r)
        if (!p->conv[r]) p->inner_iters[r] += B.h_state[p->scal0 + r].iters;
      due.push_back(p);
    }
    if (int rc = refine_all(B, due)) return rc;
  }
  if (nactive) {
    int a = 0;
    for (auto &u : B.probs) a += !u->finished;
    *nactive = a;
  }
  return LFRIC_OK;
}

}  // namespace
}  // namespace lfric

using namespace lfric;

extern "C" {

int lfric_batch_create(const lfric_solver_opts *opts, lfric_batch **out) {
  if (int rc = require_init()) return rc;
  LF_CHECK(opts && out, LFRIC_E_ARG, "null argument");
  auto *B = new lfric_batch;
  B->opts = *opts;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, ctx().device) == hipSuccess) B->ncu = prop.multiProcessorCount;
  int rc = dmalloc(reinterpret_cast<void **>(&B->d_counter), sizeof(unsigned int) * kCounters);
  if (!rc) rc = ensure_slots(*B, 8);
  // the refinement kernels use kIrBlocks partials before any table exists
  if (!rc)
    rc = ensure(reinterpret_cast<void **>(&B->d_partials), B->cap_partials, static_cast<size_t>(kIrBlocks) * 2 * kMaxR,
                sizeof(double));
  if (rc) {
    lfric_batch_destroy(B);
    return rc;
  }
  *out = B;
  return LFRIC_OK;
}

int lfric_batch_add(lfric_batch *B, const lfric_op *A64, const lfric_op *A32, const lfric_vec *diag32, lfric_vec *x,
                    const lfric_vec *b, int *id) {
  if (int rc = require_init()) return rc;
  LF_CHECK(B && A64 && A32 && diag32 && x && b && id, LFRIC_E_ARG, "null argument");
  LF_CHECK(A64->precision == LFRIC_R64 && A32->precision == LFRIC_R32 && A64->to == A32->to &&
               A64->to == A64->from && A32->to == A32->from,
           LFRIC_E_ARG, "a batch problem needs square FP64 and FP32 operators on one space");
  LF_CHECK(A32->from->ndf == 8 || A32->from->ndf == 27, LFRIC_E_UNSUPPORTED, "batched solves take ndf 8 or 27");
  LF_CHECK(x->precision == LFRIC_R64 && b->precision == LFRIC_R64 && x->nrhs == b->nrhs &&
               x->fs->undf == A64->to->undf && b->fs->undf == A64->to->undf && x->d != b->d,
           LFRIC_E_ARG, "x and b must be distinct FP64 vectors of the operator space with one number of RHS");
  LF_CHECK(diag32->precision == LFRIC_R32 && diag32->nrhs == 1 && diag32->fs->undf == A64->to->undf, LFRIC_E_ARG,
           "the Jacobi diagonal must be a one-RHS FP32 vector of the space");
  const int nr = x->nrhs;
  LF_CHECK(nr == 1 || nr == 2 || nr == 4 || nr == 8, LFRIC_E_UNSUPPORTED, "nrhs must be 1, 2, 4 or 8");
  auto p = std::make_unique<Problem>();
  p->A64 = A64;
  p->A32 = A32;
  p->diag = diag32;
  p->x = x;
  p->b = b;
  p->nrhs = nr;
  p->n = x->n;
  p->bytes32 = sizeof(float) * static_cast<size_t>(p->n) * 4;
  p->bytes64 = sizeof(double) * static_cast<size_t>(p->n);
  p->scal0 = -1;
  if (int rc = dmalloc(reinterpret_cast<void **>(&p->d), p->bytes32)) return rc;
  if (int rc = dmalloc(reinterpret_cast<void **>(&p->r64), p->bytes64)) {
    free_problem(*B, *p);
    return rc;
  }
  if (int rc = alloc_state(*B, nr, &p->scal0)) {
    free_problem(*B, *p);
    return rc;
  }
  p->r = p->d + p->n;
  p->p = p->r + p->n;
  p->q = p->p + p->n;
  p->nr0.assign(nr, 0.0);
  p->nr.assign(nr, 0.0);
  p->conv.assign(nr, 0);
  p->inner_iters.assign(nr, 0);
  // the gather lists of the space (built once per space)
  if (!A32->to->d_gcol)
    if (int rc = launch_lma_apply_mrhs(A64, x->d, p->r64, nr, LFRIC_MRHS_GATHER)) {
      free_problem(*B, *p);
      return rc;
    }
  p->id = B->next_id++;
  *id = p->id;
  B->probs.push_back(std::move(p));
  return LFRIC_OK;
}

int lfric_batch_run(lfric_batch *B, int rounds, int *nactive) {
  if (int rc = require_init()) return rc;
  LF_CHECK(B && rounds >= 0, LFRIC_E_ARG, "bad arguments");
  return run_rounds(*B, rounds, nactive);
}

int lfric_batch_solve(lfric_batch *B) {
  if (int rc = require_init()) return rc;
  LF_CHECK(B, LFRIC_E_ARG, "null argument");
  int nactive = 1;
  while (nactive > 0)
    if (int rc = run_rounds(*B, 1, &nactive)) return rc;
  return LFRIC_OK;
}

int lfric_batch_finished(const lfric_batch *B, int id, int *finished) {
  LF_CHECK(B && finished, LFRIC_E_ARG, "null argument");
  for (auto &u : B->probs)
    if (u->id == id) {
      *finished = u->finished ? 1 : 0;
      return LFRIC_OK;
    }
  return fail(LFRIC_E_ARG, "no such problem in the batch");
}

int lfric_batch_result(const lfric_batch *B, int id, lfric_solver_result *res) {
  LF_CHECK(B && res, LFRIC_E_ARG, "null argument");
  for (auto &u : B->probs)
    if (u->id == id) {
      const Problem &p = *u;
      for (int r = 0; r < p.nrhs; ++r) {
        res[r].status = p.finished ? (p.conv[r] ? LFRIC_OK : p.status == LFRIC_NAN ? LFRIC_NAN : LFRIC_NOT_CONVERGED)
                                   : LFRIC_NOT_CONVERGED;
        res[r].iterations = p.inner_iters[r];
        res[r].outer_iterations = p.outer;
        res[r].r_nrm_0 = p.nr0[r];
        res[r].r_nrm = p.nr[r];
        res[r].true_rel_res = p.nr0[r] > 0 ? p.nr[r] / p.nr0[r] : 0.0;
      }
      return LFRIC_OK;
    }
  return fail(LFRIC_E_ARG, "no such problem in the batch");
}

int lfric_batch_remove(lfric_batch *B, int id) {
  if (int rc = require_init()) return rc;
  LF_CHECK(B, LFRIC_E_ARG, "null argument");
  for (size_t i = 0; i < B->probs.size(); ++i)
    if (B->probs[i]->id == id) {
      if (B->probs[i]->inner_running) B->dirty = true;
      LF_HIP(hipStreamSynchronize(cur_stream()));
      free_problem(*B, *B->probs[i]);
      B->probs.erase(B->probs.begin() + static_cast<long>(i));
      return LFRIC_OK;
    }
  return fail(LFRIC_E_ARG, "no such problem in the batch");
}

// Autotuner: time the persistent apply of the current active set for each block size and
// grid size, keep the fastest.  Needs a round to have run (so the active set exists).
int lfric_batch_tune(lfric_batch *B, double *best_ms) {
  if (int rc = require_init()) return rc;
  LF_CHECK(B, LFRIC_E_ARG, "null argument");
  if (B->dirty)
    if (int rc = rebuild(*B)) return rc;
  if (B->active.empty()) return fail(LFRIC_E_STATE, "no active problem to tune on");
  const int blocks[] = {64, 128, 256};
  const int mults[] = {1, 2, 4, 8};
  double best = 1e30;
  int bb = B->block, bm = B->grid_mult;
  for (int blk : blocks) {
    B->block = blk;
    if (int rc = rebuild(*B)) return rc;  // tiles follow the block size
    for (int gm : mults) {
      launch_apply(*B, blk, gm);  // warm
      LF_HIP(hipEventRecord(ctx().t0, cur_stream()));
      for (int rep = 0; rep < 5; ++rep) launch_apply(*B, blk, gm);
      LF_HIP(hipEventRecord(ctx().t1, cur_stream()));
      LF_HIP(hipEventSynchronize(ctx().t1));
      float ms = 0.f;
      LF_HIP(hipEventElapsedTime(&ms, ctx().t0, ctx().t1));
      if (ms / 5 < best) {
        best = ms / 5;
        bb = blk;
        bm = gm;
      }
    }
  }
  B->block = bb;
  B->grid_mult = bm;
  if (int rc = rebuild(*B)) return rc;
  if (best_ms) *best_ms = best;
  return LFRIC_OK;
}

int lfric_batch_config(lfric_batch *B, int use_graph, int block, int grid_mult) {
  LF_CHECK(B, LFRIC_E_ARG, "null argument");
  LF_CHECK(block == 0 || block == 64 || block == 128 || block == 256, LFRIC_E_ARG, "block must be 64, 128 or 256");
  if (use_graph >= 0) B->use_graph = use_graph;
  if (block) B->block = block;
  if (grid_mult > 0) B->grid_mult = grid_mult;
  B->dirty = true;
  return LFRIC_OK;
}

int lfric_batch_set_apply(lfric_batch *B, int form) {
  LF_CHECK(B && form >= 0 && form <= 2, LFRIC_E_ARG, "bad arguments");
  B->use_colour = form;
  B->dirty = true;
  return LFRIC_OK;
}

int lfric_batch_stats(const lfric_batch *B, long long *iterations, int *block, int *grid_mult, int *nactive_tiles) {
  LF_CHECK(B, LFRIC_E_ARG, "null argument");
  if (iterations) *iterations = B->iterations;
  if (block) *block = B->block;
  if (grid_mult) *grid_mult = B->grid_mult;
  if (nactive_tiles) *nactive_tiles = B->ntiles;
  return LFRIC_OK;
}

void lfric_batch_destroy(lfric_batch *B) {
  if (!B) return;
  (void)hipStreamSynchronize(cur_stream());
  for (auto &u : B->probs) free_problem(*B, *u);
  drop_graph(*B);
  dfree(B->d_state, B->cap_state * sizeof(RhsDev));
  if (B->h_state) (void)hipHostFree(B->h_state);
  dfree(B->d_desc, B->cap_desc * sizeof(ProbDev));
  dfree(B->d_tile0, B->cap_tile0 * sizeof(int));
  dfree(B->d_vtile0, B->cap_vtile0 * sizeof(int));
  dfree(B->d_partials, B->cap_partials * sizeof(double));
  dfree(B->d_scratch, sizeof(double) * kSlot * B->scratch_slots);
  dfree(B->d_counter, sizeof(unsigned int) * kCounters);
  dfree(B->d_ctile0, B->cap_ctile0 * sizeof(int));
  if (B->h_scratch) (void)hipHostFree(B->h_scratch);
  delete B;
}

}  // extern "C"
